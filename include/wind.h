/*
 * wind_amd C-ABI: batched stiff-ODE integration (semi-implicit Euler and
 * RK2 step doubling with per-system timestep control) on AMD Instinct
 * MI355X (gfx950).
 *
 * Two layers:
 *
 * 1. Drop-in entry points of agurvich/WIND, exported with C++ linkage so the
 *    mangled names match what WIND's Python frontend looks up through ctypes
 *    (python/wind_harness.py:18-54).  They live in per-network shim libraries
 *    wind_amd/lib/<network>/{sie,rk2,sie_host}.so, mirroring WIND's
 *    cuda/lib/{sie,rk2,sie_host}.so rebuilt per network:
 *
 *      int cudaIntegrateSystem(float tnow, float tend, int n_integration_steps,
 *                              float *constants, float *equations, int Nsystems,
 *                              int Nequations_per_system, float ABSOLUTE,
 *                              float RELATIVE);
 *          symbol _Z19cudaIntegrateSystemffiPfS_iiff in sie.so and rk2.so.
 *          Replaces cuda/common/harness.cu:5-106 (one launch of the fused
 *          kernel cuda/common/kernel.cu:140-310).  Results follow the CPU
 *          c_baseline (cuda/c_baseline/common/common_gold.c:8-153), bit for
 *          bit.  Returns the total sub-step count (3 per attempt, summed over
 *          systems), truncated to int like the reference's int counter.
 *
 *      int cudaIntegrateSIE(<same 9 arguments>);
 *          symbol _Z16cudaIntegrateSIEffiPfS_iiff in sie_host.so.  Replaces
 *          cuda/SIE_host/harness_kernel.cu:427-532 (lockstep SIE: one
 *          timestep for all systems).  Returns nsteps * Nsystems.
 *
 *      void initializeCublasExternally(void);
 *          symbol _Z26initializeCublasExternallyv in sie_host.so.  Replaces
 *          the symbol the reference's harness calls before cudaIntegrateSIE
 *          (python/wind_harness.py:27,139); here it initialises the device.
 *
 * 2. The extended C API below (extern "C"): explicit networks, device-resident
 *    state, per-system step counts and status, 64-bit counters, error text.
 */
#ifndef WIND_AMD_WIND_H
#define WIND_AMD_WIND_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes returned by the extended API. */
enum wind_error {
    WIND_OK = 0,
    WIND_ERR_ARGUMENT = -1,   /* invalid argument; see wind_last_error() */
    WIND_ERR_DEVICE = -2,     /* HIP runtime error */
    WIND_ERR_UNSUPPORTED = -3 /* network / N / kernel combination not built */
};

enum wind_method { WIND_SIE = 0, WIND_RK2 = 1 };

/* Arithmetic mode.  EXACT reproduces the reference c_baseline's float
 * arithmetic operation for operation (explicit Gauss-Jordan inverse, no
 * fused multiply-add, double-precision relative error test), so results
 * are bit-identical to it.  FAST solves (I - hJ) x = f by LU elimination
 * with fused multiply-adds and hardware exp/log; results differ from the
 * reference by rounding (and, through step acceptance, by trajectory). */
enum wind_mode { WIND_MODE_EXACT = 0, WIND_MODE_FAST = 1 };

/* Timestep controller.  REFERENCE is c_baseline's (common_gold.c:116-148):
 * halve on failure (at most 10 times in a row, then accept anyway); after
 * acceptance SIE retries the whole remaining interval and RK2 doubles the
 * step; SIE keeps the two-half-step result, RK2 the extrapolation 2 y2 - y1.
 * ADAPTIVE keeps the same step-doubling pair but controls the step from the
 * error norm err = max_i |y2_i - y1_i| / (ABSOLUTE + RELATIVE |y2_i|)
 * (evaluated in double): accept iff err <= 1 (a non-finite norm fails),
 * keep the extrapolation 2 y2 - y1, and scale the step by
 * clamp(0.9 / sqrt(err), 0.2, 4). */
enum wind_controller { WIND_CTRL_REFERENCE = 0, WIND_CTRL_ADAPTIVE = 1 };

/* Linear solver of the fast mode for table-driven (mass-action) networks:
 * (I - hJ) is factored by a blocked LU in LDS whose trailing updates run on
 * the matrix cores in the chosen precision (VALU: vector FMAs, no MFMA;
 * BF16: BF16 inputs with FP32 accumulation; BF16X3: split BF16 operands,
 * hi*hi + hi*lo + lo*hi).  ir_sweeps iterative-refinement sweeps follow,
 * with the residual evaluated in double from the sparse Jacobian. */
enum wind_solver {
    WIND_SOLVER_AUTO = 0,
    WIND_SOLVER_VALU = 1,
    WIND_SOLVER_MFMA_F32 = 2,
    WIND_SOLVER_MFMA_BF16 = 3,
    WIND_SOLVER_MFMA_BF16X3 = 4
};

/* Kernel family.  AUTO picks by network and N. */
enum wind_kernel {
    WIND_KERNEL_AUTO = 0,
    WIND_KERNEL_BLOCK = 1,  /* one workgroup per system, equation per thread
                               (the reference's launch shape) */
    WIND_KERNEL_LANE = 2    /* one system per lane, persistent work queue */
};

/* Per-system termination status (extension; the reference loops forever
 * where STALLED or EXHAUSTED is reported). */
enum wind_status {
    WIND_STATUS_OK = 0,
    WIND_STATUS_STALLED = 1,   /* an accepted step did not advance tnow and
                                  the next trial would repeat it exactly,
                                  or the step underflowed to zero */
    WIND_STATUS_EXHAUSTED = 2  /* max_attempts reached */
};

/* Built-in networks. */
enum wind_builtin {
    WIND_NET_KATZ96 = 0,      /* N = 5, 10 constants per system */
    WIND_NET_NR_TEST = 1,     /* N = 2, 4 constants per system */
    WIND_NET_HHE6 = 2,        /* N = 6, 1 constant (n_H); 24 network params */
    WIND_NET_MASS_ACTION = 3, /* N species, R reactions a+b->c+d; 1 constant */
    WIND_NET_PLUGIN = 4       /* generated by wind_amd/codegen.py; see
                                 wind_network_create_plugin */
};

typedef struct wind_network wind_network;

/* Creates a handle for a compiled-in network.  `params` is required for
 * WIND_NET_HHE6 (24 floats: lnA[7], b[7], C[7], Gamma[3]) and ignored
 * otherwise. */
int wind_network_create_builtin(int builtin, const float *params, int nparams,
                                wind_network **out);

/* Loads a network compiled by the code generator (wind_amd/codegen.py)
 * from its plugin library (wind_amd/lib/gen/<name>/libwind_net.so): the
 * network's RHS and Jacobian as straight-line code, integrated by the
 * lane-per-system kernels like the compiled-in networks.  Replaces the
 * reference's network precompiler (python/ode_systems/ode_base.py:16-123),
 * which rewrites the CUDA sources and rebuilds the libraries per network.
 * WIND_ERR_ARGUMENT if the file cannot be loaded or was built against
 * another version of this library. */
int wind_network_create_plugin(const char *path, wind_network **out);

/* Constants per system of a network (its NUM_CONST). */
int wind_network_nconst(const wind_network *net);

/* Creates a mass-action network: reaction r is rx[4r] + rx[4r+1] ->
 * rx[4r+2] + rx[4r+3] with rate k[r] * c[0] * y_a * y_b.  The library copies
 * the table and builds its device-side gather lists. */
int wind_network_create_mass_action(int N, int R, const int32_t *rx, const float *k,
                                    wind_network **out);

void wind_network_destroy(wind_network *net);
int wind_network_size(const wind_network *net); /* N */

/* Per-attempt trace of wind_integrate (extension).  The reference prints
 * the same quantities for one system in its DEBUGBLOCK builds (the full-step
 * and two-half-step states of every attempt: cuda/common/kernel.cu:214-249,
 * cuda/c_baseline/common/common_gold.c:102-113).  Attempt k of system s is
 * at s * capacity + k; states at (s * capacity + k) * N.
 *   Record (optional): y1 / y2 receive each attempt's full-step and
 *   two-half-step states, flag the kernel's own error test (1 = its
 *   controller rejects the attempt), for the first `capacity` attempts.
 *   Replay (h, accept and count together): attempt k takes the step h[...]
 *   (the reference's timestep after its fmin clamp, common_gold.c:77) and
 *   is accepted iff accept[...] != 0, in place of the controller's choice;
 *   system s stops after count[s] <= capacity attempts.  An accepted attempt
 *   advances tnow by h and keeps what the controller would keep (reference
 *   SIE: y2; otherwise 2 y2 - y1).  With the oracle's own sequence the
 *   kernel integrates the same accepted-step sequence as the oracle.
 *   y_start (optional, with replay): attempt k starts from the state
 *   y_start[(s * capacity + k) * N ...] instead of the kernel's own, so that
 *   each attempt can be compared with a reference's from the same state.
 * Pointers are host or device like the other arrays (device_pointers).
 * Supported by the lane kernels of the compiled-in networks (exact and fast
 * mode) and by the dense solver of mass-action networks (fast mode);
 * other kernels return WIND_ERR_UNSUPPORTED. */
typedef struct wind_trace {
    int capacity;            /* attempts per system in each array (> 0) */
    const float *h;          /* replay: [Nsystems * capacity] */
    const int32_t *accept;   /* replay: [Nsystems * capacity] */
    const int32_t *count;    /* replay: [Nsystems] */
    const float *y_start;    /* replay, optional: [Nsystems * capacity * N] */
    float *y1;               /* record: [Nsystems * capacity * N] */
    float *y2;               /* record: [Nsystems * capacity * N] */
    int32_t *flag;           /* record: [Nsystems * capacity] */
} wind_trace;

typedef struct wind_options {
    int method;            /* wind_method */
    int mode;              /* wind_mode */
    int kernel;            /* wind_kernel */
    int const_stride;      /* floats per system in `constants`; 0 = network default */
    int device_pointers;   /* 1: constants, equations, steps_out, status_out are
                              device pointers (state stays resident) */
    int64_t max_attempts;  /* per-system attempt cap; 0 = default (2^24) */
    int32_t *steps_out;    /* optional [Nsystems]: 3 * attempts per system */
    int32_t *status_out;   /* optional [Nsystems]: wind_status */
    int64_t *total_steps;  /* optional: total sub-steps, 64-bit */
    void *stream;          /* hipStream_t; NULL = the library's stream */
    float *kernel_ms;      /* optional: device time of the integration kernel */
    int controller;        /* wind_controller */
    int solver;            /* wind_solver (fast mode, mass-action networks) */
    int ir_sweeps;         /* iterative-refinement sweeps after the solve (0-4) */
    int flags;             /* WIND_FLAG_* */
    const int32_t *cost_hint; /* wind_integrate, kernels with a work queue (the
                              lane kernels of compiled-in and generated
                              networks, the dense solver of mass-action
                              networks in fast mode): optional [Nsystems]
                              per-system cost estimates, device or host as
                              the other arrays (e.g. the previous call's
                              steps_out for the same cells).  Systems start in
                              decreasing order of it, so that long ones do not
                              finish alone at the end of the run.  Results do
                              not depend on it. */
    const wind_trace *trace;  /* wind_integrate: optional per-attempt trace or
                              replay (see wind_trace); NULL = none */
} wind_options;

/* wind_options.flags */
enum wind_flags {
    WIND_FLAG_GRAPH = 1,       /* batch: capture the fork/join launch into a hipGraph and replay it */
    WIND_FLAG_NO_AUTOTUNE = 2, /* batch: use default kernel shapes instead of tuning each bucket */
    WIND_FLAG_PIVOT = 4,       /* fast mode: partial pivoting in the LU solves (the row with the
                                  largest magnitude in the pivot column; a wave-shuffle search in
                                  the one-workgroup-per-system kernels, selects in the
                                  lane-per-system ones).  Mass-action networks then use the LDS
                                  LU instead of the blocked MFMA solver.  Exact mode rejects it:
                                  the reference's Gauss-Jordan is unpivoted. */
    WIND_FLAG_DENSE_LDS = 8    /* fast mode, mass-action SIE with 49 <= N <= 160: keep the matrix in
                                  LDS (one system per CU at N = 157) instead of the default
                                  register-resident solver (three or four systems per CU); results agree
                                  to rounding */
};

void wind_options_init(wind_options *opts);

/* Integrates every system from tnow to tend (in place in `equations`),
 * with the initial trial step (tend - tnow) / n_integration_steps.
 * Synchronous: returns after results are in `equations`. */
int wind_integrate(const wind_network *net, float tnow, float tend, int n_integration_steps,
                   const float *constants, float *equations, int Nsystems, int Neqn,
                   float ABSOLUTE, float RELATIVE, const wind_options *opts);

/* Allocation-free variant of wind_integrate.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and results as wind_integrate, with
 * these limits:
 *   - opts->device_pointers must be 1 (constants, equations, steps_out,
 *     status_out and cost_hint are device pointers; the caller owns them);
 *   - Nsystems <= WIND_NOALLOC_MAX_SYSTEMS, else WIND_ERR_ARGUMENT;
 *   - opts->kernel_ms must be NULL (WIND_ERR_ARGUMENT);
 *   - mass-action networks whose matrices do not fit in LDS (exact mode,
 *     large N) return WIND_ERR_UNSUPPORTED.
 * It calls no hipMalloc/hipFree: the work-queue counters and start-order
 * buffers live in one static device array of about 12.6 MB per GPU, held
 * until the process ends, which every call on that GPU shares.  Calls within
 * a process are serialised; the call returns after the stream (opts->stream,
 * NULL = the null stream) has finished the integration. */
#define WIND_NOALLOC_MAX_SYSTEMS (1 << 20)
int wind_integrate_noalloc(const wind_network *net, float tnow, float tend, int n_integration_steps,
                           const float *constants, float *equations, int Nsystems, int Neqn,
                           float ABSOLUTE, float RELATIVE, const wind_options *opts);

/* Output-step driver: the reference harness's outer loop
 * (python/ode_systems/ode_base.py:315-397, runIntegratorOutput) with the
 * state resident on the device between output steps.  Integrates over
 * n_output_steps intervals of dt = (tend - tnow) / n_output_steps (times
 * accumulated in double as the reference does, passed to each interval as
 * float), each with the initial trial step dt / n_integration_steps.
 * Outputs (each optional):
 *   equations_over_time  [(n_output_steps + 1) * Nsystems * Neqn] states at
 *                        the output times, the initial state first (host, or
 *                        device with opts->device_pointers);
 *   times                [n_output_steps + 1];
 *   nsteps               [n_output_steps] sub-steps per interval (64-bit);
 *   step_ms              [n_output_steps] device time per interval.
 * `equations` receives the final state.  opts->steps_out / status_out are
 * not used; opts->total_steps receives the sum of nsteps. */
int wind_integrate_outputs(const wind_network *net, double tnow, double tend, int n_output_steps,
                           int n_integration_steps, const float *constants, float *equations, int Nsystems,
                           int Neqn, float ABSOLUTE, float RELATIVE, const wind_options *opts,
                           float *equations_over_time, double *times, int64_t *nsteps, double *step_ms);

/* One bucket of a mixed batch: systems of one network, state resident on
 * the device (device pointers). */
typedef struct wind_bucket {
    const wind_network *net;
    const float *constants;   /* device: (nsys - 1) * const_stride + nconst floats */
    float *equations;         /* device: nsys * N floats, integrated in place */
    int nsys;
    int const_stride;         /* 0 = network default */
    int32_t *steps_out;       /* optional device [nsys] */
    int32_t *status_out;      /* optional device [nsys] */
    const int32_t *cost_hint; /* optional device [nsys]: per-system cost
                                 estimates, as wind_options.cost_hint for
                                 wind_integrate (e.g. the previous batch's
                                 steps_out); buckets of kernels with a work
                                 queue start their systems in decreasing
                                 order of it.  Results do not depend on it. */
} wind_bucket;

/* Integrates every bucket from tnow to tend in one fork/join launch: one
 * persistent kernel per bucket on its own stream, each with a full-device
 * grid, launched in decreasing cost per system (the dispatcher backfills the
 * CUs of buckets that finish), each bucket's kernel shape
 * (workgroup size, solver precision) chosen by an autotuner that times the
 * candidates on a sample of the bucket's systems (cached per network and
 * settings).  opts->method/mode/controller/solver apply to all buckets;
 * opts->total_steps and opts->kernel_ms report the whole batch. */
int wind_integrate_batch(const wind_bucket *buckets, int nbuckets, float tnow, float tend,
                         int n_integration_steps, float ABSOLUTE, float RELATIVE,
                         const wind_options *opts);

/* ---- Request server (continuous batching) ----
 * A server owns `nslots` execution slots (streams) and a device arena of
 * `hbm_budget_bytes`.  Submitted requests are admitted in arrival order when
 * a slot is free and their footprint (constants + state) fits in the arena,
 * then copied in, integrated (each kernel has a full-device persistent grid;
 * admitted requests overlap as earlier ones drain), and copied out.  A resident block
 * keeps a cell block's state in the arena across requests. */
typedef struct wind_server wind_server;
typedef struct wind_block wind_block;

typedef struct wind_request_desc {
    const wind_network *net;  /* transient request: network, host constants, host state in/out */
    const float *constants;
    int const_stride;         /* 0 = network default */
    float *equations;         /* host; receives the result (optional for block requests) */
    int nsys;
    wind_block *block;        /* non-NULL: integrate this resident block in place */
    float tnow, tend;
    int n_integration_steps;
    float ABSOLUTE, RELATIVE;
    int method, mode, controller;
} wind_request_desc;

int wind_server_create(size_t hbm_budget_bytes, int nslots, wind_server **out);
void wind_server_destroy(wind_server *srv);   /* drains queued and running requests */
int wind_server_block_create(wind_server *srv, const wind_network *net, const float *constants,
                             int const_stride, const float *equations, int nsys, wind_block **out);
int wind_server_block_read(wind_server *srv, wind_block *block, float *equations); /* after its requests */
void wind_server_block_destroy(wind_server *srv, wind_block *block);
int wind_server_submit(wind_server *srv, const wind_request_desc *desc, int64_t *request_id);
/* Blocks until the request is done; latency from submission to results on the host. */
int wind_server_wait(wind_server *srv, int64_t request_id, double *latency_us, int64_t *total_steps);
int wind_server_inflight(wind_server *srv, int *running, int *pending, size_t *arena_used);

/* Lockstep SIE with the semantics of the reference's cudaIntegrateSIE
 * (cuda/SIE_host/harness_kernel.cu:248-425): one shared timestep, any failing
 * system halves it for all, doubling after acceptance, return after 9
 * consecutive failures.  Host pointers.  *result = 3 * attempts * Nsystems.
 * Built for the compiled networks (not mass action). */
int wind_integrate_lockstep(const wind_network *net, float tnow, float tend,
                            int n_integration_steps, const float *constants,
                            float *equations, int Nsystems, int Neqn, float ABSOLUTE,
                            float RELATIVE, int const_stride, int64_t *result);

/* Text of the last error on the calling thread ("" if none). */
const char *wind_last_error(void);

/* Library version string. */
const char *wind_version(void);

#ifdef __cplusplus
}
#endif
#endif
