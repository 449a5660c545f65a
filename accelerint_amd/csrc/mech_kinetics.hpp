// Synthetic constant-pressure mass-action kinetics, dydt + analytic dense
// Jacobian, for the lane-group layout.
//
// This stands in for the pyJac-generated callbacks the reference expects in
// mechanism/ (examples/pyJac/pyJac.dox:9-33; callback contract
// examples/van_der_pol/dydt.cu:31-39, jacob.cu:31-43): state y = (T, Y_1..Y_ns),
// pr = pressure [Pa], Jacobian column-major J[i + j*N] = d f_i / d y_j, here
// delivered as the rows each lane owns.
//
// Model (all lanes of a group cooperate on one condition):
//   S = sum Y_k/W_k, Ctot = 1e-6 P/(R T) [mol/cm^3], C_k = Ctot (Y_k/W_k)/S
//   kf = exp(lnA + beta ln(T/1000) - EaR/T), ln Kc = -sum nu g/RT + dnu ln C0
//   q_j = kf prod C(reactants) - kf/Kc prod C(products)
//   dY_k/dt = W_k wdot_k S/Ctot,  dT/dt = -(sum H_k wdot_k) S/(Ctot cp)
// evaluated at T clamped to [kKinTmin, kKinTmax], S >= kKinSmin, cp >= kKinCpmin.
// Work split: species rows are owned as usual (row 0 = T, row k+1 = species
// k); reactions are strided over the group's lanes; per-group LDS holds the
// concentrations, the species thermo, the per-reaction rate data and, while
// the Jacobian is assembled, the concentration Jacobian d wdot / d C.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "radau_core.hpp"

namespace acc {

constexpr int kKinMaxReactions = 4096;
constexpr double kKinR = 8.31446261815324;
constexpr double kKinPatm = 101325.0;
constexpr double kKinLnTref = 6.907755278982137;  // ln(1000)
// Guards that keep f finite for any finite state: Newton trial states of a
// violent transient can carry T <= 0 or T ~ 1e7 K (ln T undefined, exp
// overflow) or mass fractions whose weighted sums vanish, and the Radau error
// norm maps NaN to its 1e-10 floor (fmax), i.e. would accept such a state.
// Physical trajectories never come near the guards.
constexpr double kKinTmin = 200.0;
constexpr double kKinTmax = 1.0e4;
constexpr double kKinSmin = 1.0e-3;
constexpr double kKinCpmin = 1.0;

struct KinParams {
  int nr;
  const double* W;       // [ns]
  const double* invW;    // [ns]
  const double* thermo;  // [ns*5]
  const int* slots;      // [nr*6]
  const double* rc;      // [nr*3] lnA, beta, EaR
  const int* csr_off;    // [ns+1] species -> (reaction, nu) list
  const int* csr_rxn;
  const int* csr_nu;
};

// Make LDS writes by some lanes of a wave visible to reads by other lanes of
// the same wave (groups never span waves).
__device__ __forceinline__ void group_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int NTOT>
struct MechKinetics {
  static constexpr int N = NTOT;
  static constexpr int NS = NTOT - 1;
  using Params = KinParams;
  static constexpr bool kJacInLds = true;

  // per group: C, gRT, hR, cpR, col [NS each]; kf, kr, q, qT [nr each];
  // Jc [NS*NS]; parked Jacobian [N*N]
  __host__ __device__ static size_t group_doubles(int nr) {
    return (size_t)5 * NS + (size_t)4 * nr + (size_t)NS * NS + (size_t)N * N;
  }
  static size_t lds_bytes(const Params& p, int groups) { return groups * group_doubles(p.nr) * sizeof(double); }

  struct Ctx {
    Params p;
    double* C;
    double* gRT;
    double* hR;
    double* cpR;
    double* col;
    double* kf;
    double* kr;
    double* q;
    double* qT;
    double* Jc;
    double* jac;
    __device__ Ctx(const Params& pp, int g, double* lds) : p(pp) {
      double* b = lds + (size_t)g * group_doubles(pp.nr);
      C = b; b += NS;
      gRT = b; b += NS;
      hR = b; b += NS;
      cpR = b; b += NS;
      col = b; b += NS;
      kf = b; b += pp.nr;
      kr = b; b += pp.nr;
      q = b; b += pp.nr;
      qT = b; b += pp.nr;
      Jc = b; b += NS * NS;
      jac = b;
    }
  };

  // Shared prologue: T, thermo of owned species (to LDS), mixture sums and
  // concentrations (to LDS).
  template <int G, int R>
  __device__ __forceinline__ static void prologue(Ctx& c, double P, const double (&y)[R], double& T,
                                                  double& S, double& Ctot, double& cpm,
                                                  double& dcpm) {
    const int ln = Grp<G>::lane();
    T = fmin(fmax(Grp<G>::bcast(y[0], 0), kKinTmin), kKinTmax);
    const double lnT = log(T);
    double Sp = 0.0, cpp = 0.0, dcp = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      if (r >= 1 && r < N) {
        const int s = r - 1;
        const double* a = c.p.thermo + 5 * s;
        const double a0 = a[0], a1 = a[1], a2 = a[2], bh = a[3], bs = a[4];
        const double cp = a0 + a1 * T + a2 * T * T;
        const double h = a0 * T + a1 * T * T / 2 + a2 * T * T * T / 3 + bh;
        const double sR = a0 * lnT + a1 * T + a2 * T * T / 2 + bs;
        c.cpR[s] = cp;
        c.hR[s] = h;
        c.gRT[s] = h / T - sR;
        const double iw = c.p.invW[s];
        Sp += y[q] * iw;
        cpp += y[q] * cp * iw;
        dcp += y[q] * (a1 + 2 * a2 * T) * iw;
      }
    }
    S = fmax(Grp<G>::sum(Sp), kKinSmin);
    cpm = fmax(kKinR * Grp<G>::sum(cpp), kKinCpmin);
    dcpm = kKinR * Grp<G>::sum(dcp);
    Ctot = 1e-6 * P / (kKinR * T);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      if (r >= 1 && r < N) {
        const int s = r - 1;
        c.C[s] = Ctot * (y[q] * c.p.invW[s]) / S;
      }
    }
    group_lds_sync();
  }

  // Rate constants and rates of progress of the reactions this lane handles.
  template <int G, bool JAC>
  __device__ __forceinline__ static void reactions(Ctx& c, double T) {
    const int ln = Grp<G>::lane();
    const double lnT = log(T);
    const double lnC0 = log(1e-6 * kKinPatm / (kKinR * T));
    const double iT = 1.0 / T;
    for (int j = ln; j < c.p.nr; j += G) {
      const int* sl = c.p.slots + 6 * j;
      const double lnA = c.p.rc[3 * j], beta = c.p.rc[3 * j + 1], EaR = c.p.rc[3 * j + 2];
      double sum_nu_g = 0.0, sum_nu_h = 0.0;
      int dnu = 0;
      double pf = 1.0, pr = 1.0;
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        const int s = sl[a];
        if (s >= 0) {
          const double sg = a < 3 ? -1.0 : 1.0;
          sum_nu_g += sg * c.gRT[s];
          if (JAC) sum_nu_h += sg * c.hR[s];
          dnu += a < 3 ? -1 : 1;
          if (a < 3) pf *= c.C[s];
          else pr *= c.C[s];
        }
      }
      const double lnkf = lnA + beta * (lnT - kKinLnTref) - EaR * iT;
      const double lnKc = -sum_nu_g + dnu * lnC0;
      const double kf = exp(lnkf);
      const double kr = exp(lnkf - lnKc);
      pf *= kf;
      pr *= kr;
      c.q[j] = pf - pr;
      if (JAC) {
        const double dlnkf = beta * iT + EaR * iT * iT;
        const double dlnkr = dlnkf - (sum_nu_h * iT * iT - dnu * iT);
        c.qT[j] = pf * dlnkf - pr * dlnkr;
        c.kf[j] = kf;
        c.kr[j] = kr;
      }
    }
    group_lds_sync();
  }

  template <int G, int R>
  __device__ static void dydt(Ctx& c, double, double P, const double (&y)[R], double (&f)[R]) {
    const int ln = Grp<G>::lane();
    double T, S, Ctot, cpm, dcpm;
    prologue<G, R>(c, P, y, T, S, Ctot, cpm, dcpm);
    reactions<G, false>(c, T);
    const double D = S / Ctot;
    double Qp = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      f[q] = 0.0;
      if (r >= 1 && r < N) {
        const int s = r - 1;
        double w = 0.0;
        const int e1 = c.p.csr_off[s + 1];
        for (int e = c.p.csr_off[s]; e < e1; ++e) w += (double)c.p.csr_nu[e] * c.q[c.p.csr_rxn[e]];
        f[q] = c.p.W[s] * w * D;
        Qp += c.hR[s] * w;
      }
    }
    const double Q = kKinR * Grp<G>::sum(Qp);
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (ln + G * q == 0) f[q] = -Q * D / cpm;
    group_lds_sync();  // LDS is reused by the next evaluation
  }

  template <int G, int R>
  __device__ static void jacob(Ctx& c, double, double P, const double (&y)[R], double (&J)[R][N]) {
    const int ln = Grp<G>::lane();
    double T, S, Ctot, cpm, dcpm;
    prologue<G, R>(c, P, y, T, S, Ctot, cpm, dcpm);
    reactions<G, true>(c, T);
    const double D = S / Ctot;
    const double iT = 1.0 / T;
    double Qp = 0.0, dQp = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      if (r >= 1 && r < N) {
        const int s = r - 1;
        double* row = c.Jc + (size_t)s * NS;
#pragma unroll
        for (int i = 0; i < NS; ++i) row[i] = 0.0;
        double w = 0.0, wT = 0.0;
        const int e1 = c.p.csr_off[s + 1];
        for (int e = c.p.csr_off[s]; e < e1; ++e) {
          const int j = c.p.csr_rxn[e];
          const double nu = (double)c.p.csr_nu[e];
          w += nu * c.q[j];
          wT += nu * c.qT[j];
          const int* sl = c.p.slots + 6 * j;
          int sp[6];
#pragma unroll
          for (int a = 0; a < 6; ++a) sp[a] = sl[a];
          double cv[6];
#pragma unroll
          for (int a = 0; a < 6; ++a) cv[a] = sp[a] >= 0 ? c.C[sp[a]] : 1.0;
          const double kfj = nu * c.kf[j], krj = -nu * c.kr[j];
#pragma unroll
          for (int b = 0; b < 6; ++b) {
            if (sp[b] >= 0) {
              double v = b < 3 ? kfj : krj;
              const int lo = b < 3 ? 0 : 3;
#pragma unroll
              for (int a = lo; a < lo + 3; ++a)
                if (a != b) v *= cv[a];
              row[sp[b]] += v;
            }
          }
        }
        // own row back into registers; r_s = sum_m Jc[s][m] C_m
        double jr[NS];
        double rs = 0.0;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          jr[i] = row[i];
          rs += jr[i] * c.C[i];
        }
        const double Ws = c.p.W[s];
        const double corr = (rs - w) / Ctot;
#pragma unroll
        for (int i = 0; i < NS; ++i) J[q][1 + i] = Ws * c.p.invW[i] * (jr[i] - corr);
        J[q][0] = Ws * D * (wT + (w - rs) * iT);
        const double hs = c.hR[s];
#pragma unroll
        for (int i = 0; i < NS; ++i) row[i] = hs * (Ctot * jr[i] - rs);
        Qp += hs * w;
        dQp += c.cpR[s] * w + hs * (wT - rs * iT);
      }
    }
    const double Q = kKinR * Grp<G>::sum(Qp);
    const double dQdT = kKinR * Grp<G>::sum(dQp);
    group_lds_sync();
    // column sums of the H-weighted species rows: lane owning row i+1 sums column i
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      if (r >= 1 && r < N) {
        const int i = r - 1;
        double s = 0.0;
        for (int k = 0; k < NS; ++k) s += c.Jc[(size_t)k * NS + i];
        c.col[i] = s;
      }
    }
    group_lds_sync();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (ln + G * q == 0) {
        const double QD = Q * D;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          const double iw = c.p.invW[i];
          const double dQdYi = kKinR * c.col[i] * iw / S;
          J[q][1 + i] = -(D * dQdYi + Q * iw / Ctot) / cpm + QD * kKinR * c.cpR[i] * iw / (cpm * cpm);
        }
        J[q][0] = -(D * dQdT + QD * iT) / cpm + QD * dcpm / (cpm * cpm);
      } else if (ln + G * q >= N) {
#pragma unroll
        for (int i = 0; i < N; ++i) J[q][i] = 0.0;
      }
    }
    group_lds_sync();
  }
};

using MechKinetics14 = MechKinetics<14>;
using MechKinetics54 = MechKinetics<54>;

// Host-side owner of a kinetics network: validation, the species -> reaction
// incidence list, and the device copies.
struct KineticsHost {
  int ns = 0, nr = 0;
  std::vector<double> W, invW, thermo, rc;
  std::vector<int> slots, csr_off, csr_rxn, csr_nu;
  void* dmem = nullptr;
  KinParams dev{};

  bool set(int ns_, int nr_, const double* W_, const double* th, const int* sl, const double* lnA,
           const double* beta, const double* EaR, std::string* msg) {
    if (ns_ <= 0 || nr_ <= 0 || !W_ || !th || !sl || !lnA || !beta || !EaR) {
      *msg = "accelerInt_set_kinetics: bad arguments";
      return false;
    }
    for (int s = 0; s < ns_; ++s)
      if (!(W_[s] > 0)) {
        *msg = "accelerInt_set_kinetics: molecular weights must be positive";
        return false;
      }
    for (int k = 0; k < 6 * nr_; ++k)
      if (sl[k] < -1 || sl[k] >= ns_) {
        *msg = "accelerInt_set_kinetics: slot species index out of range";
        return false;
      }
    ns = ns_;
    nr = nr_;
    W.assign(W_, W_ + ns);
    invW.resize(ns);
    for (int s = 0; s < ns; ++s) invW[s] = 1.0 / W[s];
    thermo.assign(th, th + 5 * ns);
    slots.assign(sl, sl + 6 * nr);
    rc.resize(3 * nr);
    for (int j = 0; j < nr; ++j) {
      rc[3 * j] = lnA[j];
      rc[3 * j + 1] = beta[j];
      rc[3 * j + 2] = EaR[j];
    }
    // net stoichiometric coefficients nu[s][j], kept in reaction order
    csr_off.assign(ns + 1, 0);
    csr_rxn.clear();
    csr_nu.clear();
    std::vector<std::vector<std::pair<int, int>>> lists(ns);
    for (int j = 0; j < nr; ++j) {
      int nu[6] = {0};
      int spc[6];
      int m = 0;
      for (int a = 0; a < 6; ++a) {
        const int s = slots[6 * j + a];
        if (s < 0) continue;
        int idx = -1;
        for (int b = 0; b < m; ++b)
          if (spc[b] == s) idx = b;
        if (idx < 0) {
          idx = m++;
          spc[idx] = s;
        }
        nu[idx] += a < 3 ? -1 : 1;
      }
      for (int b = 0; b < m; ++b)
        if (nu[b] != 0) lists[spc[b]].push_back({j, nu[b]});
    }
    for (int s = 0; s < ns; ++s) {
      csr_off[s + 1] = csr_off[s] + (int)lists[s].size();
      for (auto& e : lists[s]) {
        csr_rxn.push_back(e.first);
        csr_nu.push_back(e.second);
      }
    }
    return true;
  }

  bool upload(std::string* msg) {
    release();
    const size_t nW = ns, nT = 5 * (size_t)ns, nS = 6 * (size_t)nr, nR = 3 * (size_t)nr;
    const size_t nO = ns + 1, nE = csr_rxn.size() ? csr_rxn.size() : 1;
    const size_t bytes = (2 * nW + nT + nR) * sizeof(double) + (nS + nO + 2 * nE) * sizeof(int) + 256;
    if (hipMalloc(&dmem, bytes) != hipSuccess) {
      *msg = "hipMalloc failed for kinetics tables";
      dmem = nullptr;
      return false;
    }
    char* p = (char*)dmem;
    auto put = [&](const void* src, size_t n, size_t esz) {
      void* dst = p;
      if (n) hipMemcpy(dst, src, n * esz, hipMemcpyHostToDevice);
      p += ((n * esz + 15) / 16) * 16;
      return dst;
    };
    dev.nr = nr;
    dev.W = (const double*)put(W.data(), nW, sizeof(double));
    dev.invW = (const double*)put(invW.data(), nW, sizeof(double));
    dev.thermo = (const double*)put(thermo.data(), nT, sizeof(double));
    dev.rc = (const double*)put(rc.data(), nR, sizeof(double));
    dev.slots = (const int*)put(slots.data(), nS, sizeof(int));
    dev.csr_off = (const int*)put(csr_off.data(), nO, sizeof(int));
    dev.csr_rxn = (const int*)put(csr_rxn.data(), csr_rxn.size(), sizeof(int));
    dev.csr_nu = (const int*)put(csr_nu.data(), csr_nu.size(), sizeof(int));
    return hipDeviceSynchronize() == hipSuccess;
  }

  template <int NTOT>
  KinParams params() const {
    return dev;
  }

  void release() {
    if (dmem) hipFree(dmem);
    dmem = nullptr;
  }
};

}  // namespace acc
