/*
 * TEST INFRASTRUCTURE ONLY - the oracle.
 *
 * CPU restatement of agurvich/WIND's c_baseline integrators (SIE and RK2
 * step doubling with the unpivoted Gauss-Jordan inverse) and of the ODE
 * networks the tests and benchmarks use.  It is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in wind_amd/ links or calls it.
 *
 * Reference files restated (paths relative to the WIND tree):
 *   cuda/c_baseline/common/common_gold.c:8-153      checkError, integrateSystem
 *   cuda/c_baseline/sie/sie_gold.c:10-119           acceptSolution, scaleAndInvertJacobians,
 *                                                   take_step, goldIntegrateSystem (SIE)
 *   cuda/c_baseline/rk2/rk2_gold.c:8-79             acceptSolution, take_step,
 *                                                   goldIntegrateSystem (RK2)
 *   cuda/c_baseline/linear_algebra/gje_gold.c:5-144 setIdentity, scaleRow, subtractRows,
 *                                                   gjeUFactor, gjeLFactor, gjeInvertMatrix
 *   cuda/c_baseline/linear_algebra/matrix_ops.c:6-27 matrixVectorMult
 *   cuda/ode_system/ode_gold.c:3-93                 Katz96 calculate_dydt / calculate_jacobian
 *   python/ode_systems/NR_test.py:106-121           NR_test generated RHS / Jacobian text
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   gje_invert, matvec, check_error, take_step, integrate_system, gold_integrate:
 *       bit-exact against the unmodified reference c_baseline compiled from
 *       source (oracle/ref.mk -> oracle/_ref/), through committed golden vectors
 *       (tests/golden/), plus closed forms (2x2 inverse, A*inv(A) = I,
 *       J = 0 -> explicit Euler, diagonal J, NR_test analytic solution).
 *   katz96 network: bit-exact against the reference's compiled ode_gold.c.
 *   nr_test network: exact linear closed form (J constant, f = J y).
 *   hhe6, mass_action networks (authored here, not in the reference):
 *       float64 finite-difference Jacobian, element conservation (sum H,
 *       sum He; sum y for A+B->C+D), independent numpy float64 RHS.
 *   wexpf / wlogf: against libm expf/logf to <= 2 ulp over their domain.
 */
#ifndef WIND_ORACLE_H
#define WIND_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum oracle_network {
    ORACLE_NET_KATZ96 = 0,      /* N = 5, 10 constants (ode_gold.c) */
    ORACLE_NET_NR_TEST = 1,     /* N = 2, 4 constants (NR_test.py) */
    ORACLE_NET_HHE6 = 2,        /* N = 6, 1 constant: n_H */
    ORACLE_NET_MASS_ACTION = 3, /* table driven, 1 constant: rate scale */
    ORACLE_NET_GENERATED = 4    /* functions written by wind_amd/codegen.py */
};

enum oracle_method { ORACLE_SIE = 0, ORACLE_RK2 = 1 };

/* Timestep controller: REFERENCE restates common_gold.c:116-148; ADAPTIVE is
 * the library's error-norm controller (an extension, see wind_oracle.cpp). */
enum oracle_controller { ORACLE_CTRL_REFERENCE = 0, ORACLE_CTRL_ADAPTIVE = 1 };

/* Per-system termination status (the reference has no such guards). */
enum oracle_status {
    ORACLE_STATUS_OK = 0,
    ORACLE_STATUS_STALLED = 1,    /* accepted step made no progress; would repeat forever */
    ORACLE_STATUS_EXHAUSTED = 2   /* max_attempts reached */
};

/* Network-level parameters.  hhe6 uses the first 24 floats of `params`
 * (lnA[7], b[7], C[7], Gamma[3]); mass_action uses the reaction table. */
typedef struct oracle_net {
    int id;
    int N;
    const float *params;      /* hhe6 */
    int R;                    /* mass_action reactions */
    const int32_t *rx;        /* [R*4]: reactant a, reactant b, product c, product d */
    const float *k;           /* [R] rate constants */
    /* generated networks: RHS and column-major Jacobian (J accumulated onto
     * zeros) in float and double, from oracle/_build/gen/<name>.so */
    void (*gen_rhs_f32)(const float *y, const float *c, float *f);
    void (*gen_jac_f32)(const float *y, const float *c, float *J);
    void (*gen_rhs_f64)(const double *y, const double *c, double *f);
    void (*gen_jac_f64)(const double *y, const double *c, double *J);
} oracle_net;

/* One attempt of the step-doubling loop, recorded for trace fixtures. */
typedef struct oracle_trace {
    int capacity;             /* attempts the buffers can hold */
    int count;                /* attempts recorded (may exceed capacity) */
    float *tnow;              /* [capacity] */
    float *h;                 /* [capacity] */
    int32_t *flag;            /* [capacity] error flag before the cap test */
    float *y1;                /* [capacity*N] */
    float *y2;                /* [capacity*N] */
    int32_t *accept;          /* [capacity] 1 = the attempt was accepted (optional) */
} oracle_trace;

/* Restates goldIntegrateSystem (sie_gold.c:81-119 / rk2_gold.c:48-79) with
 * the constant stride and network passed in rather than compiled in.
 * `steps_out` (optional) gets 3 * attempts per system. Returns the sum. */
int oracle_gold_integrate(int method, const oracle_net *net,
                          float tnow, float tend, int n_integration_steps,
                          const float *constants, int const_stride,
                          float *equations, int Nsystems, int Neqn,
                          float ABSOLUTE, float RELATIVE,
                          int32_t *steps_out);

/* As above, with an attempt cap (0 = none) and per-system status. */
int oracle_gold_integrate_ex(int method, const oracle_net *net,
                             float tnow, float tend, int n_integration_steps,
                             const float *constants, int const_stride,
                             float *equations, int Nsystems, int Neqn,
                             float ABSOLUTE, float RELATIVE,
                             int32_t *steps_out, long long max_attempts,
                             int32_t *status_out, int controller);

/* Same algorithm in double (state, time and tolerances), used for
 * FP64 reference integrations. */
int oracle_gold_integrate_f64(int method, const oracle_net *net,
                              double tnow, double tend, int n_integration_steps,
                              const double *constants, int const_stride,
                              double *equations, int Nsystems, int Neqn,
                              double ABSOLUTE, double RELATIVE,
                              int32_t *steps_out, int controller);

/* One system with an attempt trace (restates integrateSystem,
 * common_gold.c:48-153, and records what its DEBUGBLOCK build prints,
 * common_gold.c:102-113: y1 and y2 of every attempt). */
int oracle_integrate_system_traced(int method, const oracle_net *net,
                                   float tnow, float tend, float timestep,
                                   float *y, const float *c, int N,
                                   float ABSOLUTE, float RELATIVE,
                                   oracle_trace *trace, long long max_attempts,
                                   int32_t *status, int controller);

/* oracle_gold_integrate_ex with every system's attempt trace, laid out as
 * the library's wind_trace: attempt k of system s at s*capacity + k (h,
 * accept, flag; y1/y2 at (s*capacity + k)*N), count[s] attempts recorded
 * (at most capacity; a system with more attempts is integrated to the end
 * and count[s] = -attempts).  Systems are split over nthreads threads. */
void oracle_integrate_traced_batch(int method, const oracle_net *net, float tnow, float tend,
                                   int n_integration_steps, const float *constants, int const_stride,
                                   float *equations, int Nsystems, int Neqn, float ABSOLUTE, float RELATIVE,
                                   int capacity, float *h, int32_t *accept, int32_t *flag, float *y1,
                                   float *y2, int32_t *count, int32_t *steps_out, int32_t *status_out,
                                   long long max_attempts, int controller, int nthreads);

/* Replays recorded attempt sequences (h and accept of every attempt in the
 * wind_trace layout above, count[s] attempts of system s, one output
 * interval starting at tnow) on the recorded run's float time grid: attempt
 * k runs from t_k to t_k' = fl32(t_k + h_k) with the full step t_k' - t_k
 * and two half steps of (t_k' - t_k) / 2, as integrate_system forms them,
 * and an accepted attempt moves t to t_k'.  An accepted attempt keeps
 * 2 y2 - y1 (adaptive controller, and the reference controller's RK2) or y2
 * (the reference controller's SIE).
 *   whole run (ystart null): `equations` [Nsystems*N] starts from the
 *       initial states and ends with the replayed final states;
 *   per attempt (ystart [Nsystems*capacity*N]): every attempt starts from
 *       ystart of that attempt, and y1/y2 (optional) receive its states.
 * The float version replays the float run bit for bit; the double version
 * gives the same steps in float64 (the float grid's step sizes are exact in
 * double), the exact-arithmetic value of each attempt or run. */
void oracle_replay_batch(int method, const oracle_net *net, float tnow, const float *constants, int const_stride,
                         float *equations, int Nsystems, int N, int capacity, const float *h,
                         const int32_t *accept, const int32_t *count, const float *ystart, float *y1, float *y2,
                         int controller, int nthreads);
void oracle_replay_batch_f64(int method, const oracle_net *net, float tnow, const float *constants,
                             int const_stride, double *equations, int Nsystems, int N, int capacity,
                             const float *h, const int32_t *accept, const int32_t *count, const float *ystart,
                             double *y1, double *y2, int controller, int nthreads);

/* Multi-threaded variant of oracle_gold_integrate for the CPU baseline:
 * systems are split into contiguous slices, one slice per thread. */
int oracle_gold_integrate_mt(int method, const oracle_net *net,
                             float tnow, float tend, int n_integration_steps,
                             const float *constants, int const_stride,
                             float *equations, int Nsystems, int Neqn,
                             float ABSOLUTE, float RELATIVE,
                             int32_t *steps_out, int nthreads, int controller);

/* Lockstep SIE with the semantics of cuda/SIE_host/harness_kernel.cu:248-425
 * and c_baseline sub-step arithmetic.  Returns 3 * attempts * Nsystems. */
long long oracle_lockstep(const oracle_net *net, float tnow, float tend, int n_integration_steps,
                          const float *constants, int const_stride, float *equations, int Nsystems,
                          int Neqn, float ABSOLUTE, float RELATIVE);

/* Stage-level entry points for known-answer tests. */
void oracle_gje_invert(float *matrix, float *inverse, int N);              /* gje_gold.c:136-144 */
void oracle_matvec(const float *matrix, float *vector, int N);            /* matrix_ops.c:6-27 */
int oracle_check_error(int method, const float *y1, const float *y2, int N,
                       float ABSOLUTE, float RELATIVE);                   /* common_gold.c:8-46 */
int oracle_take_step(int method, const oracle_net *net, float tnow, float tend,
                     int n, float *y, const float *c, int N);             /* sie_gold.c:38-79 */
void oracle_rhs(const oracle_net *net, float t, const float *y, const float *c, float *f);
void oracle_jac(const oracle_net *net, const float *y, const float *c, float *J); /* J zeroed first */
void oracle_rhs_f64(const oracle_net *net, double t, const double *y, const double *c, double *f);
void oracle_jac_f64(const oracle_net *net, const double *y, const double *c, double *J);
float oracle_wexpf(float x);
float oracle_wlogf(float x);

#ifdef __cplusplus
}
#endif
#endif
