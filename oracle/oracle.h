/*
 * CPU oracle for the batched Radau IIA (order 5) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it.  The product library never
 * links or calls it.
 *
 * It restates, function by function, the reference's CPU integrator
 * radau2a/radau2a.c (control constants :37-66, tableau :101-231, kernels
 * :256-508, integrate :520-736), the OpenMP driver generic/solver_generic.c:28-58,
 * and the LAPACK semantics of dgetrf_/zgetrf_/dgetrs_/zgetrs_ that the
 * reference links against (generic/lapack_dfns.h:12-15).  Tolerances and the
 * CONST_TIME_STEP switch are runtime parameters here instead of the
 * reference's generated compile-time macros.
 */
#ifndef ACCELERINT_ORACLE_H
#define ACCELERINT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest system size the oracle accepts (stack arrays are sized by n). */
#define ORA_MAX_N 256

/* pyJac-style callbacks (examples/van_der_pol/dydt.h:24, jacob.h:25) with an
 * extra opaque context pointer standing in for the compiled-in mechanism. */
typedef void (*ora_dydt_fn)(double t, double pr, const double *y, double *dy, const void *ctx);
typedef void (*ora_jac_fn)(double t, double pr, const double *y, double *jac, const void *ctx);

typedef struct {
    int n;                 /* NSP */
    double atol;           /* ATOL */
    double rtol;           /* RTOL */
    int const_time_step;   /* CONST_TIME_STEP */
    ora_dydt_fn dydt;
    ora_jac_fn jac;
    const void *ctx;
    int fd_jac;            /* 1: eval_jacob is generic/fd_jacob.c over dydt */
    int fd_order;          /* FD_ORD of generic/fd_jacob.c:14: 1, 2, 4 or 6 */
} ora_problem;

/* OR-ed into a batch driver's mechanism kind: use the finite-difference
 * Jacobian (generic/fd_jacob.c) instead of the analytic one. */
#define ORA_FD_JACOBIAN 256
/* Finite-difference order (1, 2, 4, 6) in bits 9-11 of the kind; 0 means 1. */
#define ORA_FD_ORDER_SHIFT 9

/* Per-condition event counters (algorithmic work accounting). */
enum {
    ORA_CNT_STEP = 0,     /* Nsteps increments (attempted steps past the LU) */
    ORA_CNT_ACCEPT,       /* accepted steps */
    ORA_CNT_REJECT,       /* error-test rejections */
    ORA_CNT_NEWTON_FAIL,  /* Newton non-convergence restarts */
    ORA_CNT_LU,           /* RK_Decomp calls (E1 and, unless singular, E2) */
    ORA_CNT_LU_FAIL,      /* RK_Decomp calls that reported info != 0 */
    ORA_CNT_NEWTON,       /* Newton iterations (RK_PrepareRHS + RK_Solve) */
    ORA_CNT_F,            /* dydt calls */
    ORA_CNT_JAC,          /* eval_jacob calls */
    ORA_CNT_ES,           /* real back-solves inside RK_ErrorEstimate */
    ORA_NCNT
};

/* Return codes: radau2a/radau2a_props.h:30-39 */
#define ORA_EC_SUCCESS 0
#define ORA_EC_CONSECUTIVE_STEPS 1
#define ORA_EC_MAX_STEPS_EXCEEDED 2
#define ORA_EC_H_PLUS_T_EQUALS_H 3
#define ORA_EC_NEWTON_MAX_ITERATIONS_EXCEEDED 4

/* LAPACK-semantics dense LU (column-major, 1-based pivots, info>0 on the
 * first exactly-zero pivot). */
void ora_dgetrf(int n, double *a, int *ipiv, int *info);
void ora_zgetrf(int n, double *a /* interleaved re,im */, int *ipiv, int *info);
void ora_dgetrs(int n, const double *a, const int *ipiv, double *b);
void ora_zgetrs(int n, const double *a, const int *ipiv, double *b /* interleaved */);

/* Tableau constants (radau2a.c:101-231) in the order A, c, e, gamma, alpha,
 * beta, T, TinvAinv, ELO (45 doubles). */
void ora_tableau(double *out);

/* radau2a.c:520 integrate(); counters may be NULL. */
int ora_integrate(const ora_problem *p, double t_start, double t_end, double pr, double *y,
                  int64_t *counters);

/* generic/solver_generic.c:28 intDriver(): y_global is column-major
 * y[tid + i*num]; codes[num] and counters[num*ORA_NCNT] may be NULL.
 * nthreads <= 0 keeps the OpenMP default. */
void ora_int_driver(const ora_problem *p, int num, double t, double t_end, const double *pr_global,
                    double *y_global, int *codes, int64_t *counters, int nthreads);

/* ---- built-in callback sets ------------------------------------------------ */

/* van der Pol (examples/van_der_pol/dydt.c:28-35, jacob.c:29-41); pr = mu. */
void ora_vdp_dydt(double t, double pr, const double *y, double *dy, const void *ctx);
void ora_vdp_jac(double t, double pr, const double *y, double *jac, const void *ctx);

/* Linear constant-coefficient system y' = pr * M y (M column-major n x n,
 * ctx points at ora_linear).  Used to pin the method against closed forms. */
typedef struct { int n; const double *m; } ora_linear;
void ora_linear_dydt(double t, double pr, const double *y, double *dy, const void *ctx);
void ora_linear_jac(double t, double pr, const double *y, double *jac, const void *ctx);

/* Synthetic constant-pressure mass-action kinetics (ctx points at ora_kinetics).
 * State y = (T, Y_1..Y_ns); pr = pressure [Pa]. */
typedef struct {
    int ns;               /* species */
    int nr;               /* reactions */
    const double *W;      /* [ns] molecular weight, kg/mol */
    const double *thermo; /* [ns*5] a0, a1, a2, bh, bs: cp/R = a0 + a1 T + a2 T^2,
                             h/R = a0 T + a1 T^2/2 + a2 T^3/3 + bh,
                             s/R = a0 ln T + a1 T + a2 T^2/2 + bs */
    const int *slots;     /* [nr*6] reactant slots 0..2, product slots 3..5, -1 empty */
    const double *lnA;    /* [nr] ln of pre-exponential (mol, cm, s units) at 1000 K */
    const double *beta;   /* [nr] temperature exponent of (T / 1000 K) */
    const double *EaR;    /* [nr] activation temperature Ea/R, K */
} ora_kinetics;
void ora_kin_dydt(double t, double pr, const double *y, double *dy, const void *ctx);
void ora_kin_jac(double t, double pr, const double *y, double *jac, const void *ctx);
/* Per-reaction net rates of progress q_j [mol cm^-3 s^-1]. */
void ora_kin_rates(const ora_kinetics *k, double T, double P, const double *Y, double *q);

/* Batch drivers for ctypes callers: mech 0 = van der Pol, 1 = linear, 2 = kinetics. */
void ora_batch(int mech, const void *ctx, int n, double atol, double rtol, int const_time_step,
               int num, double t, double t_end, const double *pr, double *y_colmajor,
               int *codes, int64_t *counters, int nthreads);
void ora_eval(int mech, const void *ctx, int n, int what, double t, double pr, const double *y,
              double *out);
/* Finite-difference Jacobian, generic/fd_jacob.c:24-127 (FD_ORD from the
 * kind's order bits). */
void ora_fd_jacob(int mech, const void *ctx, int n, double atol, double rtol, double t, double pr,
                  const double *y, double *jac);
void ora_problem_fd_jacob(const ora_problem *p, double t, double pr, const double *cy, double *jac);

#ifdef __cplusplus
}
#endif
#endif
