/*
 * CPU oracle callbacks: van der Pol, linear test systems and the synthetic
 * constant-pressure kinetics; batch entry points for ctypes.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

/* ---- van der Pol: examples/van_der_pol/dydt.c:28-35, jacob.c:29-41 -------- */
void ora_vdp_dydt(double t, double mu, const double *y, double *dy, const void *ctx) {
    (void)t; (void)ctx;
    dy[0] = y[1];
    dy[1] = mu * (1 - y[0] * y[0]) * y[1] - y[0];
}

void ora_vdp_jac(double t, double mu, const double *y, double *jac, const void *ctx) {
    (void)t; (void)ctx;
    jac[0] = 0;                             /* d f0 / d y0 */
    jac[1] = -2 * mu * y[0] * y[1] - 1;     /* d f1 / d y0 */
    jac[2] = 1;                             /* d f0 / d y1 */
    jac[3] = mu * (1 - y[0] * y[0]);        /* d f1 / d y1 */
}

/* ---- linear y' = pr * M y ------------------------------------------------- */
void ora_linear_dydt(double t, double pr, const double *y, double *dy, const void *ctx) {
    const ora_linear *L = (const ora_linear *)ctx;
    int n = L->n;
    (void)t;
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += L->m[i + j * n] * y[j];
        dy[i] = pr * s;
    }
}

void ora_linear_jac(double t, double pr, const double *y, double *jac, const void *ctx) {
    const ora_linear *L = (const ora_linear *)ctx;
    int n = L->n;
    (void)t; (void)y;
    for (int k = 0; k < n * n; ++k) jac[k] = pr * L->m[k];
}

/* ---- synthetic constant-pressure kinetics --------------------------------
 *
 * State y = (T, Y_1..Y_ns), pr = P [Pa].  Concentrations in mol/cm^3:
 *   S = sum Y_k / W_k,  Ctot = 1e-6 P / (R T),  C_k = Ctot (Y_k/W_k) / S.
 * Rates of progress (mass action, reversible):
 *   ln kf = lnA + beta ln(T/1000) - EaR/T
 *   ln Kc = -sum_k nu_k g_k/(RT) + dnu ln C0,  C0 = 1e-6 P_atm/(R T)
 *   q = kf prod_{reactant slots} C - kf/Kc prod_{product slots} C
 * Right-hand side:
 *   dY_k/dt = W_k wdot_k S / Ctot,   dT/dt = -(sum_k H_k wdot_k) S / (Ctot cp)
 * with H_k = R (h_k/R) molar enthalpy and cp = R sum_k Y_k (cp_k/R) / W_k.
 * Everything is evaluated at T clamped to [KIN_TMIN, KIN_TMAX], with
 * S >= KIN_SMIN and cp >= KIN_CPMIN.
 */
#define KIN_R 8.31446261815324
#define KIN_PATM 101325.0
#define KIN_TREF 1000.0
/* Guards applied inside the callbacks so that f stays finite for any finite
 * state: Newton trial states of a violent transient can carry T <= 0 or
 * T ~ 1e7 K (ln T undefined, exp overflow) or mass fractions whose weighted
 * sums vanish.  The reference's RK_ErrorNorm maps NaN to 1e-10
 * (fmax(sqrt(NaN), 1e-10), radau2a.c:452), so a NaN trial state would be
 * accepted as converged.  Physical trajectories never come near the guards. */
#define KIN_TMIN 200.0
#define KIN_TMAX 1.0e4
#define KIN_SMIN 1.0e-3
#define KIN_CPMIN 1.0

typedef struct {
    double S, Ctot, cpm, dcpm_dT;
} kin_mix;

static void kin_thermo(const ora_kinetics *k, double T, double *cp_R, double *h_R, double *g_RT) {
    double lnT = log(T);
    for (int s = 0; s < k->ns; ++s) {
        const double *a = k->thermo + 5 * s;
        cp_R[s] = a[0] + a[1] * T + a[2] * T * T;
        h_R[s] = a[0] * T + a[1] * T * T / 2 + a[2] * T * T * T / 3 + a[3];
        double s_R = a[0] * lnT + a[1] * T + a[2] * T * T / 2 + a[4];
        g_RT[s] = h_R[s] / T - s_R;
    }
}

/* forward/reverse rate constants of reaction j and their T-log-derivatives */
static void kin_rate_consts(const ora_kinetics *k, int j, double T, double lnC0, const double *h_R,
                            const double *g_RT, double *kf, double *kr, double *dlnkf,
                            double *dlnkr) {
    const int *sl = k->slots + 6 * j;
    double sum_nu_g = 0, sum_nu_h = 0;
    int dnu = 0;
    for (int a = 0; a < 6; ++a) {
        int s = sl[a];
        if (s < 0) continue;
        double sg = a < 3 ? -1.0 : 1.0;
        sum_nu_g += sg * g_RT[s];
        sum_nu_h += sg * h_R[s];
        dnu += a < 3 ? -1 : 1;
    }
    double lnkf = k->lnA[j] + k->beta[j] * log(T / KIN_TREF) - k->EaR[j] / T;
    double lnKc = -sum_nu_g + dnu * lnC0;
    *kf = exp(lnkf);
    *kr = exp(lnkf - lnKc);
    *dlnkf = k->beta[j] / T + k->EaR[j] / (T * T);
    /* d lnKc/dT = sum nu h/(R T^2) - dnu / T */
    *dlnkr = *dlnkf - (sum_nu_h / (T * T) - dnu / T);
}

static void kin_conc(const ora_kinetics *k, double T, double P, const double *Y, double *C,
                     kin_mix *mx) {
    double S = 0;
    for (int s = 0; s < k->ns; ++s) S += Y[s] / k->W[s];
    S = fmax(S, KIN_SMIN);
    double Ctot = 1e-6 * P / (KIN_R * T);
    for (int s = 0; s < k->ns; ++s) C[s] = Ctot * (Y[s] / k->W[s]) / S;
    mx->S = S;
    mx->Ctot = Ctot;
}

void ora_kin_rates(const ora_kinetics *k, double T, double P, const double *Y, double *q) {
    int ns = k->ns;
    double C[ns], cp_R[ns], h_R[ns], g_RT[ns];
    kin_mix mx;
    kin_conc(k, T, P, Y, C, &mx);
    kin_thermo(k, T, cp_R, h_R, g_RT);
    double lnC0 = log(1e-6 * KIN_PATM / (KIN_R * T));
    for (int j = 0; j < k->nr; ++j) {
        double kf, kr, d1, d2;
        kin_rate_consts(k, j, T, lnC0, h_R, g_RT, &kf, &kr, &d1, &d2);
        const int *sl = k->slots + 6 * j;
        double pf = kf, pr = kr;
        for (int a = 0; a < 3; ++a) if (sl[a] >= 0) pf *= C[sl[a]];
        for (int a = 3; a < 6; ++a) if (sl[a] >= 0) pr *= C[sl[a]];
        q[j] = pf - pr;
    }
}

void ora_kin_dydt(double t, double P, const double *y, double *dy, const void *ctx) {
    const ora_kinetics *k = (const ora_kinetics *)ctx;
    int ns = k->ns;
    (void)t;
    double T = fmin(fmax(y[0], KIN_TMIN), KIN_TMAX);
    const double *Y = y + 1;
    double q[k->nr], wdot[ns], cp_R[ns], h_R[ns], g_RT[ns], C[ns];
    kin_mix mx;
    kin_conc(k, T, P, Y, C, &mx);
    kin_thermo(k, T, cp_R, h_R, g_RT);
    double lnC0 = log(1e-6 * KIN_PATM / (KIN_R * T));
    for (int s = 0; s < ns; ++s) wdot[s] = 0;
    for (int j = 0; j < k->nr; ++j) {
        double kf, kr, d1, d2;
        kin_rate_consts(k, j, T, lnC0, h_R, g_RT, &kf, &kr, &d1, &d2);
        const int *sl = k->slots + 6 * j;
        double pf = kf, pr = kr;
        for (int a = 0; a < 3; ++a) if (sl[a] >= 0) pf *= C[sl[a]];
        for (int a = 3; a < 6; ++a) if (sl[a] >= 0) pr *= C[sl[a]];
        q[j] = pf - pr;
        for (int a = 0; a < 6; ++a) {
            if (sl[a] < 0) continue;
            wdot[sl[a]] += a < 3 ? -q[j] : q[j];
        }
    }
    double cpm = 0, Q = 0;
    for (int s = 0; s < ns; ++s) {
        cpm += Y[s] * cp_R[s] / k->W[s];
        Q += h_R[s] * wdot[s];
    }
    cpm = fmax(cpm * KIN_R, KIN_CPMIN);
    Q *= KIN_R;
    double D = mx.S / mx.Ctot;
    dy[0] = -Q * D / cpm;
    for (int s = 0; s < ns; ++s) dy[1 + s] = k->W[s] * wdot[s] * D;
}

/* Analytic dense Jacobian, column-major jac[i + j*N], N = ns + 1.
 * With Jc = d wdot / d C (at fixed T), r_k = sum_m Jc[k][m] C_m and
 * D = S / Ctot:
 *   dC_m/dY_i = (Ctot delta_mi - C_m) / (W_i S),  dC_m/dT = -C_m / T
 *   J[Yk][Yi] = (W_k / W_i) (Jc[k][i] - (r_k - wdot_k) / Ctot)
 *   J[Yk][T]  = W_k D (wdotT_k + (wdot_k - r_k) / T)
 *   J[T][Yi]  = -(D dQ/dY_i + Q/(W_i Ctot)) / cp + Q D R cp_i/R / (W_i cp^2)
 *   J[T][T]   = -(D dQ/dT + Q D / T) / cp + Q D dcp/dT / cp^2
 * where dQ/dY_i = sum_k H_k (Ctot Jc[k][i] - r_k)/(W_i S) and
 * dQ/dT = sum_k R cp_k/R wdot_k + H_k (wdotT_k - r_k / T). */
void ora_kin_jac(double t, double P, const double *y, double *jac, const void *ctx) {
    const ora_kinetics *k = (const ora_kinetics *)ctx;
    const int ns = k->ns, N = ns + 1;
    (void)t;
    double T = fmin(fmax(y[0], KIN_TMIN), KIN_TMAX);
    const double *Y = y + 1;
    double wdot[ns], wdotT[ns], cp_R[ns], h_R[ns], g_RT[ns], C[ns], r[ns];
    double Jc[ns * ns]; /* row-major Jc[kk*ns + m] */
    kin_mix mx;
    kin_conc(k, T, P, Y, C, &mx);
    kin_thermo(k, T, cp_R, h_R, g_RT);
    double lnC0 = log(1e-6 * KIN_PATM / (KIN_R * T));
    memset(wdot, 0, sizeof(wdot));
    memset(wdotT, 0, sizeof(wdotT));
    memset(Jc, 0, sizeof(Jc));
    for (int j = 0; j < k->nr; ++j) {
        double kf, kr, dlnkf, dlnkr;
        kin_rate_consts(k, j, T, lnC0, h_R, g_RT, &kf, &kr, &dlnkf, &dlnkr);
        const int *sl = k->slots + 6 * j;
        double pf = kf, pr = kr;
        for (int a = 0; a < 3; ++a) if (sl[a] >= 0) pf *= C[sl[a]];
        for (int a = 3; a < 6; ++a) if (sl[a] >= 0) pr *= C[sl[a]];
        double q = pf - pr;
        double qT = pf * dlnkf - pr * dlnkr;
        /* dq/dC at each slot b: product of the other factors of its side */
        double dq[6];
        for (int b = 0; b < 6; ++b) {
            dq[b] = 0;
            if (sl[b] < 0) continue;
            int lo = b < 3 ? 0 : 3;
            double v = b < 3 ? kf : -kr;
            for (int a = lo; a < lo + 3; ++a) {
                if (a == b || sl[a] < 0) continue;
                v *= C[sl[a]];
            }
            dq[b] = v;
        }
        for (int a = 0; a < 6; ++a) {
            if (sl[a] < 0) continue;
            double sg = a < 3 ? -1.0 : 1.0;
            wdot[sl[a]] += sg * q;
            wdotT[sl[a]] += sg * qT;
            for (int b = 0; b < 6; ++b) {
                if (sl[b] < 0) continue;
                Jc[sl[a] * ns + sl[b]] += sg * dq[b];
            }
        }
    }
    for (int kk = 0; kk < ns; ++kk) {
        double s = 0;
        for (int m = 0; m < ns; ++m) s += Jc[kk * ns + m] * C[m];
        r[kk] = s;
    }
    double cpm = 0, dcpm = 0, Q = 0, dQdT = 0;
    for (int s = 0; s < ns; ++s) {
        const double *a = k->thermo + 5 * s;
        cpm += Y[s] * cp_R[s] / k->W[s];
        dcpm += Y[s] * (a[1] + 2 * a[2] * T) / k->W[s];
        Q += h_R[s] * wdot[s];
        dQdT += cp_R[s] * wdot[s] + h_R[s] * (wdotT[s] - r[s] / T);
    }
    cpm = fmax(cpm * KIN_R, KIN_CPMIN); dcpm *= KIN_R; Q *= KIN_R; dQdT *= KIN_R;
    const double S = mx.S, Ctot = mx.Ctot, D = S / Ctot;
    /* species rows */
    for (int kk = 0; kk < ns; ++kk) {
        double Wk = k->W[kk];
        double corr = (r[kk] - wdot[kk]) / Ctot;
        for (int i = 0; i < ns; ++i)
            jac[(1 + kk) + (1 + i) * N] = (Wk / k->W[i]) * (Jc[kk * ns + i] - corr);
        jac[(1 + kk) + 0 * N] = Wk * D * (wdotT[kk] + (wdot[kk] - r[kk]) / T);
    }
    /* temperature row */
    for (int i = 0; i < ns; ++i) {
        double s = 0;
        for (int kk = 0; kk < ns; ++kk) s += h_R[kk] * (Ctot * Jc[kk * ns + i] - r[kk]);
        double dQdYi = KIN_R * s / (k->W[i] * S);
        jac[0 + (1 + i) * N] = -(D * dQdYi + Q / (k->W[i] * Ctot)) / cpm +
                               Q * D * KIN_R * cp_R[i] / (k->W[i] * cpm * cpm);
    }
    jac[0] = -(D * dQdT + Q * D / T) / cpm + Q * D * dcpm / (cpm * cpm);
}

/* ---- ctypes entry points -------------------------------------------------- */

static void make_problem(ora_problem *p, int mech, const void *ctx, int n, double atol,
                         double rtol, int cts) {
    p->n = n;
    p->atol = atol;
    p->rtol = rtol;
    p->const_time_step = cts;
    p->ctx = ctx;
    p->fd_jac = (mech & ORA_FD_JACOBIAN) != 0;
    p->fd_order = (mech >> ORA_FD_ORDER_SHIFT) & 7;
    if (p->fd_order == 0) p->fd_order = 1;
    mech &= 0xff;
    switch (mech) {
        case 0: p->dydt = ora_vdp_dydt; p->jac = ora_vdp_jac; break;
        case 1: p->dydt = ora_linear_dydt; p->jac = ora_linear_jac; break;
        default: p->dydt = ora_kin_dydt; p->jac = ora_kin_jac; break;
    }
}

void ora_batch(int mech, const void *ctx, int n, double atol, double rtol, int const_time_step,
               int num, double t, double t_end, const double *pr, double *y_colmajor,
               int *codes, int64_t *counters, int nthreads) {
    ora_problem p;
    make_problem(&p, mech, ctx, n, atol, rtol, const_time_step);
    ora_int_driver(&p, num, t, t_end, pr, y_colmajor, codes, counters, nthreads);
}

/* what: 0 = dydt (out n), 1 = analytic Jacobian (out n*n column-major) */
void ora_eval(int mech, const void *ctx, int n, int what, double t, double pr, const double *y,
              double *out) {
    ora_problem p;
    make_problem(&p, mech, ctx, n, 1e-10, 1e-6, 0);
    if (what == 0) p.dydt(t, pr, y, out, p.ctx);
    else p.jac(t, pr, y, out, p.ctx);
}

/* Central-difference stencils of generic/fd_jacob.c:33-66 (FD_ORD 2, 4, 6):
 * offsets in units of the increment and weights, the textbook central
 * differences (antisymmetric weights, first moment 1, higher moments zero up
 * to the order). */
static void fd_stencil(int order, int k, double *x, double *w) {
    static const double x2[2] = {-1.0, 1.0}, w2[2] = {-0.5, 0.5};
    static const double x4[4] = {-2.0, -1.0, 1.0, 2.0};
    static const double w4[4] = {1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0};
    static const double x6[6] = {-3.0, -2.0, -1.0, 1.0, 2.0, 3.0};
    static const double w6[6] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
    if (order == 2) { *x = x2[k]; *w = w2[k]; }
    else if (order == 4) { *x = x4[k]; *w = w4[k]; }
    else { *x = x6[k]; *w = w6[k]; }
}

/* generic/fd_jacob.c:24-127: FD_ORD 1 forward differences (:94-101), 2/4/6
 * central stencils summed in stencil order and divided by r (:103-119) */
void ora_problem_fd_jacob(const ora_problem *p, double t, double pr, const double *cy, double *jac) {
    const int n = p->n;
    double y[n], dy[n], ewt[n], ftemp[n];
    memcpy(y, cy, n * sizeof(double));
    p->dydt(t, pr, y, dy, p->ctx);
    for (int i = 0; i < n; ++i) ewt[i] = p->atol + (p->rtol * fabs(y[i]));
    double srur = sqrt(DBL_EPSILON);
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += (ewt[i] * dy[i]) * (ewt[i] * dy[i]);
    double fac = sqrt(sum / ((double)n));
    double r0 = 1000.0 * p->rtol * DBL_EPSILON * ((double)n) * fac;
    for (int j = 0; j < n; ++j) {
        double yj = y[j];
        double r = fmax(srur * fabs(yj), r0 / ewt[j]);
        if (p->fd_order <= 1) {
            y[j] = yj + r;
            p->dydt(t, pr, y, ftemp, p->ctx);
            for (int i = 0; i < n; ++i) jac[i + n * j] = (ftemp[i] - dy[i]) / r;
        } else {
            for (int i = 0; i < n; ++i) jac[i + n * j] = 0.0;
            for (int k = 0; k < p->fd_order; ++k) {
                double xk, wk;
                fd_stencil(p->fd_order, k, &xk, &wk);
                y[j] = yj + xk * r;
                p->dydt(t, pr, y, ftemp, p->ctx);
                for (int i = 0; i < n; ++i) jac[i + n * j] += wk * ftemp[i];
            }
            for (int i = 0; i < n; ++i) jac[i + n * j] /= r;
        }
        y[j] = yj;
    }
}

/* ctypes: the stencil of an order, for the moment checks of the tests */
void ora_fd_stencil(int order, int k, double *x, double *w) { fd_stencil(order, k, x, w); }

void ora_fd_jacob(int mech, const void *ctx, int n, double atol, double rtol, double t, double pr,
                  const double *cy, double *jac) {
    ora_problem p;
    make_problem(&p, mech, ctx, n, atol, rtol, 0);
    ora_problem_fd_jacob(&p, t, pr, cy, jac);
}
