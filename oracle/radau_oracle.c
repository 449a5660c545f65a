/*
 * CPU oracle: restatement of the reference's CPU Radau IIA-5 integrator.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the part of
 * radau2a/radau2a.c it follows.  All matrices are column-major, a[i + j*n].
 * Runtime n replaces the compile-time NSP; arrays are C99 VLAs on the stack,
 * as the reference keeps its workspaces on the stack (radau2a.c:538-556).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdbool.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- control constants: radau2a.c:37-66 ---------------------------------- */
#define Max_no_steps (200000)
#define NewtonMaxit (8)
#define StartNewton (true)
#define Roundoff (DBL_EPSILON)
#define FacMin (0.2)
#define FacMax (8)
#define FacSafe (0.9)
#define FacRej (0.1)
#define ThetaMin (0.001)
#define NewtonTol (0.03)
#define Qmin (1.0)
#define Qmax (1.2)
#define Max_consecutive_errs (5)

/* ---- tableau: radau2a.c:101-231 (classical error estimator branch) -------
 * These are the published Radau IIA(5) coefficients (Hairer & Wanner, RADAU5)
 * at 34 significant digits; the test suite re-derives every one of them from
 * closed forms. */
static const double rkA[3][3] = {
    {1.968154772236604258683861429918299e-1, -6.55354258501983881085227825696087e-2,
     2.377097434822015242040823210718965e-2},
    {3.944243147390872769974116714584975e-1, 2.920734116652284630205027458970589e-1,
     -4.154875212599793019818600988496743e-2},
    {3.764030627004672750500754423692808e-1, 5.124858261884216138388134465196080e-1,
     1.111111111111111111111111111111111e-1}};
static const double rkC[3] = {1.550510257216821901802715925294109e-1,
                              6.449489742783178098197284074705891e-1, 1.0};
static const double rkE[4] = {0.05, -10.04880939982741556246032950764708 * 0.05,
                              1.382142733160748895793662840980412 * 0.05,
                              -0.3333333333333333333333333333333333 * 0.05};
static const double rkGamma = 3.637834252744495732208418513577775;
static const double rkAlpha = 2.681082873627752133895790743211112;
static const double rkBeta = 3.050430199247410569426377624787569;
static const double rkT[3][3] = {
    {9.443876248897524148749007950641664e-2, -1.412552950209542084279903838077973e-1,
     -3.00291941051474244918611170890539e-2},
    {2.502131229653333113765090675125018e-1, 2.041293522937999319959908102983381e-1,
     3.829421127572619377954382335998733e-1},
    {1.0, 1.0, 0.0}};
static const double rkTinvAinv[3][3] = {
    {1.520148562492775501049204957366528e+1, 1.192055789400527921212348994770778,
     1.903956760517560343018332287285119},
    {-9.669512977505946748632625374449567, -8.724028436822336183071773193986487,
     3.096043239482439656981667712714881},
    {-1.409513259499574544876303981551774e+1, 5.895975725255405108079130152868952,
     -1.441236197545344702389881889085515e-1}};
static const double rkELO = 4;

/* Copies the tableau into out[0..44]: A (9, row-major), c (3), e (4),
 * gamma, alpha, beta, T (9), TinvAinv (9), ELO, so tests can check every
 * constant against its closed form. */
void ora_tableau(double *out) {
    int k = 0;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[k++] = rkA[i][j];
    for (int i = 0; i < 3; ++i) out[k++] = rkC[i];
    for (int i = 0; i < 4; ++i) out[k++] = rkE[i];
    out[k++] = rkGamma; out[k++] = rkAlpha; out[k++] = rkBeta;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[k++] = rkT[i][j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[k++] = rkTinvAinv[i][j];
    out[k++] = rkELO;
}

/* ---- LAPACK semantics ---------------------------------------------------- */

/* Smith's complex division (a+ib)/(c+id). */
static inline void cdiv(double a, double b, double c, double d, double *re, double *im) {
    if (fabs(c) >= fabs(d)) {
        double r = d / c, den = c + d * r;
        *re = (a + b * r) / den;
        *im = (b - a * r) / den;
    } else {
        double r = c / d, den = c * r + d;
        *re = (a * r + b) / den;
        *im = (b * r - a) / den;
    }
}

/* dgetf2: unblocked right-looking LU with partial pivoting (idamax = first
 * maximum of |a|); pivots 1-based; info = index of the first zero pivot. */
void ora_dgetrf(int n, double *a, int *ipiv, int *info) {
    const double sfmin = DBL_MIN;
    *info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double vmax = fabs(a[j + j * n]);
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(a[i + j * n]);
            if (v > vmax) { vmax = v; jp = i; }
        }
        ipiv[j] = jp + 1;
        if (a[jp + j * n] != 0.0) {
            if (jp != j) {
                for (int k = 0; k < n; ++k) {
                    double tmp = a[j + k * n];
                    a[j + k * n] = a[jp + k * n];
                    a[jp + k * n] = tmp;
                }
            }
            if (j < n - 1) {
                double piv = a[j + j * n];
                if (fabs(piv) >= sfmin) {
                    double r = 1.0 / piv;
                    for (int i = j + 1; i < n; ++i) a[i + j * n] *= r;
                } else {
                    for (int i = j + 1; i < n; ++i) a[i + j * n] /= piv;
                }
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
        if (j < n - 1) {  /* dger: A22 -= x * y' column by column, skipping y_k == 0 */
            for (int k = j + 1; k < n; ++k) {
                double yk = a[j + k * n];
                if (yk != 0.0) {
                    double temp = -yk;
                    for (int i = j + 1; i < n; ++i) a[i + k * n] += a[i + j * n] * temp;
                }
            }
        }
    }
}

/* zgetf2: as above for complex (interleaved re,im); izamax compares
 * |Re| + |Im| (dcabs1). */
void ora_zgetrf(int n, double *a, int *ipiv, int *info) {
    const double sfmin = DBL_MIN;
#define RE(i, j) a[2 * ((i) + (j) * n)]
#define IM(i, j) a[2 * ((i) + (j) * n) + 1]
    *info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double vmax = fabs(RE(j, j)) + fabs(IM(j, j));
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(RE(i, j)) + fabs(IM(i, j));
            if (v > vmax) { vmax = v; jp = i; }
        }
        ipiv[j] = jp + 1;
        if (RE(jp, j) != 0.0 || IM(jp, j) != 0.0) {
            if (jp != j) {
                for (int k = 0; k < n; ++k) {
                    double tr = RE(j, k), ti = IM(j, k);
                    RE(j, k) = RE(jp, k); IM(j, k) = IM(jp, k);
                    RE(jp, k) = tr; IM(jp, k) = ti;
                }
            }
            if (j < n - 1) {
                double pr_ = RE(j, j), pi_ = IM(j, j);
                if (hypot(pr_, pi_) >= sfmin) {
                    double rr, ri;
                    cdiv(1.0, 0.0, pr_, pi_, &rr, &ri);
                    for (int i = j + 1; i < n; ++i) {
                        double xr = RE(i, j), xi = IM(i, j);
                        RE(i, j) = xr * rr - xi * ri;
                        IM(i, j) = xr * ri + xi * rr;
                    }
                } else {
                    for (int i = j + 1; i < n; ++i) {
                        double qr, qi;
                        cdiv(RE(i, j), IM(i, j), pr_, pi_, &qr, &qi);
                        RE(i, j) = qr; IM(i, j) = qi;
                    }
                }
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
        if (j < n - 1) {  /* zgeru */
            for (int k = j + 1; k < n; ++k) {
                double yr = RE(j, k), yi = IM(j, k);
                if (yr != 0.0 || yi != 0.0) {
                    double tr = -yr, ti = -yi;
                    for (int i = j + 1; i < n; ++i) {
                        double xr = RE(i, j), xi = IM(i, j);
                        RE(i, k) += xr * tr - xi * ti;
                        IM(i, k) += xr * ti + xi * tr;
                    }
                }
            }
        }
    }
#undef RE
#undef IM
}

/* dgetrs('N', nrhs = 1): dlaswp, then column-oriented dtrsm (unit lower,
 * non-unit upper) skipping zero right-hand-side entries as reference BLAS does. */
void ora_dgetrs(int n, const double *a, const int *ipiv, double *b) {
    for (int i = 0; i < n; ++i) {
        int ip = ipiv[i] - 1;
        if (ip != i) { double t = b[i]; b[i] = b[ip]; b[ip] = t; }
    }
    for (int k = 0; k < n; ++k) {
        if (b[k] != 0.0) {
            for (int i = k + 1; i < n; ++i) b[i] -= b[k] * a[i + k * n];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        if (b[k] != 0.0) {
            b[k] /= a[k + k * n];
            for (int i = 0; i < k; ++i) b[i] -= b[k] * a[i + k * n];
        }
    }
}

void ora_zgetrs(int n, const double *a, const int *ipiv, double *b) {
    for (int i = 0; i < n; ++i) {
        int ip = ipiv[i] - 1;
        if (ip != i) {
            double tr = b[2 * i], ti = b[2 * i + 1];
            b[2 * i] = b[2 * ip]; b[2 * i + 1] = b[2 * ip + 1];
            b[2 * ip] = tr; b[2 * ip + 1] = ti;
        }
    }
    for (int k = 0; k < n; ++k) {
        double br = b[2 * k], bi = b[2 * k + 1];
        if (br != 0.0 || bi != 0.0) {
            for (int i = k + 1; i < n; ++i) {
                double ar = a[2 * (i + k * n)], ai = a[2 * (i + k * n) + 1];
                b[2 * i] -= br * ar - bi * ai;
                b[2 * i + 1] -= br * ai + bi * ar;
            }
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double br = b[2 * k], bi = b[2 * k + 1];
        if (br != 0.0 || bi != 0.0) {
            cdiv(br, bi, a[2 * (k + k * n)], a[2 * (k + k * n) + 1], &br, &bi);
            b[2 * k] = br; b[2 * k + 1] = bi;
            for (int i = 0; i < k; ++i) {
                double ar = a[2 * (i + k * n)], ai = a[2 * (i + k * n) + 1];
                b[2 * i] -= br * ar - bi * ai;
                b[2 * i + 1] -= br * ai + bi * ar;
            }
        }
    }
}

/* ---- integrator kernels --------------------------------------------------- */

/* radau2a.c:74-81 scale() and :88-94 scale_init() */
static void scale(int n, double atol, double rtol, const double *y0, const double *y, double *sc) {
    for (int i = 0; i < n; ++i) sc[i] = 1.0 / (atol + fmax(fabs(y0[i]), fabs(y[i])) * rtol);
}
static void scale_init(int n, double atol, double rtol, const double *y0, double *sc) {
    for (int i = 0; i < n; ++i) sc[i] = 1.0 / (atol + fabs(y0[i]) * rtol);
}

/* radau2a.c:256-278 RK_Decomp: E1 = gamma/H - J, E2 = (alpha + i beta)/H - J,
 * then dgetrf_ and (if E1 is regular) zgetrf_. */
static void RK_Decomp(int n, double H, double *E1, double *E2, const double *Jac, int *ipiv1,
                      int *ipiv2, int *info) {
    double t2r = rkAlpha / H, t2i = rkBeta / H;
    double t1 = rkGamma / H;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            E1[i + j * n] = -Jac[i + j * n];
            E2[2 * (i + j * n)] = -Jac[i + j * n];
            E2[2 * (i + j * n) + 1] = 0.0;
        }
        E1[i + i * n] += t1;
        E2[2 * (i + i * n)] += t2r;
        E2[2 * (i + i * n) + 1] += t2i;
    }
    ora_dgetrf(n, E1, ipiv1, info);
    if (*info != 0) return;
    ora_zgetrf(n, E2, ipiv2, info);
}

/* radau2a.c:283-297 RK_Make_Interpolate */
static void RK_Make_Interpolate(int n, const double *Z1, const double *Z2, const double *Z3,
                                double *CONT) {
    double den = (rkC[2] - rkC[1]) * (rkC[1] - rkC[0]) * (rkC[0] - rkC[2]);
    for (int i = 0; i < n; i++) {
        CONT[i] = ((-rkC[2] * rkC[2] * rkC[1] * Z1[i] + Z3[i] * rkC[1] * rkC[0] * rkC[0] +
                    rkC[1] * rkC[1] * rkC[2] * Z1[i] - rkC[1] * rkC[1] * rkC[0] * Z3[i] +
                    rkC[2] * rkC[2] * rkC[0] * Z2[i] - Z2[i] * rkC[2] * rkC[0] * rkC[0]) /
                   den) -
                  Z3[i];
        CONT[n + i] = -(rkC[0] * rkC[0] * (Z3[i] - Z2[i]) + rkC[1] * rkC[1] * (Z1[i] - Z3[i]) +
                        rkC[2] * rkC[2] * (Z2[i] - Z1[i])) /
                      den;
        CONT[2 * n + i] = (rkC[0] * (Z3[i] - Z2[i]) + rkC[1] * (Z1[i] - Z3[i]) +
                           rkC[2] * (Z2[i] - Z1[i])) /
                          den;
    }
}

/* radau2a.c:302-314 RK_Interpolate.  The third stage uses x2 as its outer
 * factor exactly as the reference does (:312). */
static void RK_Interpolate(int n, double H, double Hold, double *Z1, double *Z2, double *Z3,
                           const double *CONT) {
    double r = H / Hold;
    double x1 = 1.0 + rkC[0] * r;
    double x2 = 1.0 + rkC[1] * r;
    double x3 = 1.0 + rkC[2] * r;
    for (int i = 0; i < n; i++) {
        Z1[i] = CONT[i] + x1 * (CONT[n + i] + x1 * CONT[2 * n + i]);
        Z2[i] = CONT[i] + x2 * (CONT[n + i] + x2 * CONT[2 * n + i]);
        Z3[i] = CONT[i] + x2 * (CONT[n + i] + x3 * CONT[2 * n + i]);
    }
}

/* radau2a.c:353-383 RK_PrepareRHS: R_k = Z_k - H sum_j A_kj f(t + c_j H, y + Z_j) */
static void RK_PrepareRHS(const ora_problem *p, double t, double pr, double H, const double *Y,
                          const double *Z1, const double *Z2, const double *Z3, double *R1,
                          double *R2, double *R3) {
    int n = p->n;
    double TMP[n], F[n];
    const double *Z[3] = {Z1, Z2, Z3};
    for (int i = 0; i < n; i++) { R1[i] = Z1[i]; R2[i] = Z2[i]; R3[i] = Z3[i]; }
    for (int s = 0; s < 3; ++s) {
        for (int i = 0; i < n; i++) TMP[i] = Y[i] + Z[s][i];              /* WADD */
        p->dydt(t + rkC[s] * H, pr, TMP, F, p->ctx);
        double a1 = -H * rkA[0][s], a2 = -H * rkA[1][s], a3 = -H * rkA[2][s];
        for (int i = 0; i < n; i++) {                                     /* DAXPY3 */
            R1[i] += a1 * F[i];
            R2[i] += a2 * F[i];
            R3[i] += a3 * F[i];
        }
    }
}

/* radau2a.c:388-441 RK_Solve */
static void RK_Solve(int n, double H, const double *E1, const double *E2, double *R1, double *R2,
                     double *R3, const int *ipiv1, const int *ipiv2) {
    for (int i = 0; i < n; i++) {
        double x1 = R1[i] / H, x2 = R2[i] / H, x3 = R3[i] / H;
        R1[i] = rkTinvAinv[0][0] * x1 + rkTinvAinv[0][1] * x2 + rkTinvAinv[0][2] * x3;
        R2[i] = rkTinvAinv[1][0] * x1 + rkTinvAinv[1][1] * x2 + rkTinvAinv[1][2] * x3;
        R3[i] = rkTinvAinv[2][0] * x1 + rkTinvAinv[2][1] * x2 + rkTinvAinv[2][2] * x3;
    }
    ora_dgetrs(n, E1, ipiv1, R1);
    double temp[2 * n];
    for (int i = 0; i < n; ++i) { temp[2 * i] = R2[i]; temp[2 * i + 1] = R3[i]; }
    ora_zgetrs(n, E2, ipiv2, temp);
    for (int i = 0; i < n; ++i) { R2[i] = temp[2 * i]; R3[i] = temp[2 * i + 1]; }
    for (int i = 0; i < n; ++i) {
        double x1 = R1[i], x2 = R2[i], x3 = R3[i];
        R1[i] = rkT[0][0] * x1 + rkT[0][1] * x2 + rkT[0][2] * x3;
        R2[i] = rkT[1][0] * x1 + rkT[1][1] * x2 + rkT[1][2] * x3;
        R3[i] = rkT[2][0] * x1 + rkT[2][1] * x2 + rkT[2][2] * x3;
    }
}

/* radau2a.c:446-453 RK_ErrorNorm */
static double RK_ErrorNorm(int n, const double *sc, const double *DY) {
    double sum = 0;
    for (int i = 0; i < n; ++i) sum += (sc[i] * sc[i] * DY[i] * DY[i]);
    return fmax(sqrt(sum / ((double)n)), 1e-10);
}

/* radau2a.c:458-508 RK_ErrorEstimate */
static double RK_ErrorEstimate(const ora_problem *p, double H, double t, double pr,
                               const double *Y, const double *F0, const double *Z1,
                               const double *Z2, const double *Z3, const double *sc,
                               const double *E1, const int *ipiv1, bool FirstStep, bool Reject,
                               int64_t *cnt) {
    int n = p->n;
    double HrkE1 = rkE[1] / H, HrkE2 = rkE[2] / H, HrkE3 = rkE[3] / H;
    double F1[n], F2[n], TMP[n];
    for (int i = 0; i < n; ++i) F2[i] = HrkE1 * Z1[i] + HrkE2 * Z2[i] + HrkE3 * Z3[i];
    for (int i = 0; i < n; ++i) TMP[i] = rkE[0] * F0[i] + F2[i];
    ora_dgetrs(n, E1, ipiv1, TMP);
    cnt[ORA_CNT_ES]++;
    double Err = RK_ErrorNorm(n, sc, TMP);
    if (Err >= 1.0 && (FirstStep || Reject)) {
        for (int i = 0; i < n; i++) TMP[i] += Y[i];
        p->dydt(t, pr, TMP, F1, p->ctx);
        cnt[ORA_CNT_F]++;
        for (int i = 0; i < n; i++) TMP[i] = F1[i] + F2[i];
        ora_dgetrs(n, E1, ipiv1, TMP);
        cnt[ORA_CNT_ES]++;
        Err = RK_ErrorNorm(n, sc, TMP);
    }
    return Err;
}

/* radau2a.c:520-736 integrate() */
int ora_integrate(const ora_problem *p, double t_start, double t_end, double pr, double *y,
                  int64_t *counters) {
    const int n = p->n;
    const double atol = p->atol, rtol = p->rtol;
    int64_t cnt_local[ORA_NCNT];
    int64_t *cnt = counters ? counters : cnt_local;
    memset(cnt, 0, sizeof(int64_t) * ORA_NCNT);

    double Hmin = 0, Hold = 0, Hacc = 0, ErrOld = 0;
    double H = p->const_time_step ? (t_end - t_start) : fmin(5e-7, t_end - t_start);
    double Hnew;
    double t = t_start;
    bool Reject = false, FirstStep = true, SkipJac = false, SkipLU = false;
    double sc[n], A[n * n], E1[n * n], E2[2 * n * n];
    int ipiv1[n], ipiv2[n];
    double Z1[n], Z2[n], Z3[n], DZ1[n], DZ2[n], DZ3[n], CONT[3 * n], y0[n], F0[n];
    memset(A, 0, sizeof(A)); memset(E1, 0, sizeof(E1)); memset(E2, 0, sizeof(E2));
    memset(ipiv1, 0, sizeof(ipiv1)); memset(ipiv2, 0, sizeof(ipiv2));
    memset(Z1, 0, sizeof(Z1)); memset(Z2, 0, sizeof(Z2)); memset(Z3, 0, sizeof(Z3));
    memset(DZ1, 0, sizeof(DZ1)); memset(DZ2, 0, sizeof(DZ2)); memset(DZ3, 0, sizeof(DZ3));
    memset(CONT, 0, sizeof(CONT));
    scale_init(n, atol, rtol, y, sc);
    memcpy(y0, y, n * sizeof(double));
    int info = 0, Nconsecutive = 0, Nsteps = 0;
    double NewtonRate = pow(2.0, 1.25);

    while (t + Roundoff < t_end) {
        if (!Reject) {
            p->dydt(t, pr, y, F0, p->ctx);
            cnt[ORA_CNT_F]++;
        }
        if (!SkipLU) {
            if (!SkipJac) {
                if (p->fd_jac) ora_problem_fd_jacob(p, t, pr, y, A);
                else p->jac(t, pr, y, A, p->ctx);
                cnt[ORA_CNT_JAC]++;
            }
            RK_Decomp(n, H, E1, E2, A, ipiv1, ipiv2, &info);
            cnt[ORA_CNT_LU]++;
            if (info != 0) {
                cnt[ORA_CNT_LU_FAIL]++;
                Nconsecutive += 1;
                if (Nconsecutive >= Max_consecutive_errs) return ORA_EC_CONSECUTIVE_STEPS;
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
        cnt[ORA_CNT_STEP]++;
        if (Nsteps >= Max_no_steps) return ORA_EC_MAX_STEPS_EXCEEDED;
        if (0.1 * fabs(H) <= fabs(t) * Roundoff) return ORA_EC_H_PLUS_T_EQUALS_H;
        if (FirstStep || !StartNewton) {
            memset(Z1, 0, n * sizeof(double));
            memset(Z2, 0, n * sizeof(double));
            memset(Z3, 0, n * sizeof(double));
        } else {
            RK_Interpolate(n, H, Hold, Z1, Z2, Z3, CONT);
        }
        bool NewtonDone = false;
        double NewtonIncrementOld = 0;
        double Fac = 0.5;
        int NewtonIter = 0;
        double Theta = 0;
        NewtonRate = pow(fmax(NewtonRate, DBL_EPSILON), 0.8);

        for (; NewtonIter < NewtonMaxit; NewtonIter++) {
            RK_PrepareRHS(p, t, pr, H, y, Z1, Z2, Z3, DZ1, DZ2, DZ3);
            cnt[ORA_CNT_F] += 3;
            RK_Solve(n, H, E1, E2, DZ1, DZ2, DZ3, ipiv1, ipiv2);
            cnt[ORA_CNT_NEWTON]++;
            double d1 = RK_ErrorNorm(n, sc, DZ1);
            double d2 = RK_ErrorNorm(n, sc, DZ2);
            double d3 = RK_ErrorNorm(n, sc, DZ3);
            double NewtonIncrement = sqrt((d1 * d1 + d2 * d2 + d3 * d3) / 3.0);
            Theta = ThetaMin;
            if (NewtonIter > 0) {
                Theta = NewtonIncrement / NewtonIncrementOld;
                if (Theta < 0.99) {
                    NewtonRate = Theta / (1.0 - Theta);
                } else {
                    break;
                }
                if (NewtonIter < NewtonMaxit) {
                    double NewtonPredictedErr =
                        (NewtonIncrement * pow(Theta, (NewtonMaxit - NewtonIter - 1))) /
                        (1.0 - Theta);
                    if (NewtonPredictedErr >= NewtonTol) {
                        double Qnewton = fmin(10.0, NewtonPredictedErr / NewtonTol);
                        Fac = 0.8 * pow(Qnewton, -1.0 / ((double)(NewtonMaxit - NewtonIter)));
                        break;
                    }
                }
            }
            NewtonIncrementOld = fmax(NewtonIncrement, Roundoff);
            for (int i = 0; i < n; i++) {
                Z1[i] -= DZ1[i];
                Z2[i] -= DZ2[i];
                Z3[i] -= DZ3[i];
            }
            NewtonDone = (NewtonRate * NewtonIncrement <= NewtonTol);
            if (NewtonDone) break;
            if (NewtonIter == NewtonMaxit - 1) return ORA_EC_NEWTON_MAX_ITERATIONS_EXCEEDED;
        }
        if (p->const_time_step) {
            /* radau2a.c:725-733: constant stepping, no error control */
            t += H;
            for (int i = 0; i < n; i++) y[i] += Z3[i];
            cnt[ORA_CNT_ACCEPT]++;
            continue;
        }
        if (!NewtonDone) {
            cnt[ORA_CNT_NEWTON_FAIL]++;
            H = Fac * H;
            Reject = true;
            SkipJac = true;
            SkipLU = false;
            continue;
        }
        double Err = RK_ErrorEstimate(p, H, t, pr, y, F0, Z1, Z2, Z3, sc, E1, ipiv1, FirstStep,
                                      Reject, cnt);
        Fac = pow(Err, (-1.0 / rkELO)) * (1.0 + 2 * NewtonMaxit) /
              (NewtonIter + 1.0 + 2 * NewtonMaxit);
        Fac = fmin(FacMax, fmax(FacMin, Fac));
        Hnew = Fac * H;
        if (Err < 1.0) {
            cnt[ORA_CNT_ACCEPT]++;
            if (!FirstStep) {  /* Gustafsson predictive control, :676-685 */
                double FacGus = FacSafe * (H / Hacc) * pow(Err * Err / ErrOld, -0.25);
                FacGus = fmin(FacMax, fmax(FacMin, FacGus));
                Fac = fmin(Fac, FacGus);
                Hnew = Fac * H;
            }
            Hacc = H;
            ErrOld = fmax(1e-2, Err);
            FirstStep = false;
            Hold = H;
            t += H;
            for (int i = 0; i < n; i++) y[i] += Z3[i];
            if (StartNewton) RK_Make_Interpolate(n, Z1, Z2, Z3, CONT);
            scale(n, atol, rtol, y, y0, sc);
            memcpy(y0, y, n * sizeof(double));
            Hnew = fmin(fmax(Hnew, Hmin), t_end - t);
            if (Reject) Hnew = fmin(Hnew, H);
            Reject = false;
            if (t + Hnew / Qmin - t_end >= 0.0) {
                H = t_end - t;
            } else {
                double Hratio = Hnew / H;
                SkipLU = (Theta <= ThetaMin) && (Hratio >= Qmin) && (Hratio <= Qmax);
                if (!SkipLU) H = Hnew;
            }
            SkipJac = NewtonIter == 1 || NewtonRate <= ThetaMin;
        } else {
            cnt[ORA_CNT_REJECT]++;
            if (FirstStep || Reject) {
                H = FacRej * H;
            } else {
                H = Hnew;
            }
            Reject = true;
            SkipJac = true;
            SkipLU = false;
        }
    }
    return ORA_EC_SUCCESS;
}

/* generic/solver_generic.c:28-58 intDriver(): OpenMP static schedule over
 * conditions; gather/scatter column-major y[tid + i*num].  The reference
 * aborts the process on the first non-zero code (check_error); the oracle
 * returns the codes instead. */
void ora_int_driver(const ora_problem *p, int num, double t, double t_end, const double *pr_global,
                    double *y_global, int *codes, int64_t *counters, int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int tid = 0; tid < num; ++tid) {
        int n = p->n;
        double y_local[n];
        for (int i = 0; i < n; i++) y_local[i] = y_global[tid + (size_t)i * num];
        int code = ora_integrate(p, t, t_end, pr_global[tid], y_local,
                                 counters ? counters + (size_t)tid * ORA_NCNT : NULL);
        if (codes) codes[tid] = code;
        for (int i = 0; i < n; i++) y_global[tid + (size_t)i * num] = y_local[i];
    }
    (void)nthreads;
}
