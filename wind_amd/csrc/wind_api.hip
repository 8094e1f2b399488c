// Host side of the wind_amd C-ABI (include/wind.h): device context and
// buffer pool, network handles, argument checks, kernel dispatch.
//
// Replaces the per-call host harness of the reference
// (cuda/common/harness.cu:5-106), which allocates, copies, launches, copies
// back and frees on every call.  Here device buffers are pooled per process
// (grow-only, guarded by a mutex) and reused across calls, and callers that
// keep their state on the device pass device pointers and skip the copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels_small.h"
#include "wind_internal.h"

using namespace wind;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define WIND_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(WIND_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Pool {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Context {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Pool constants, equations, steps, status, counters;
};

Context &ctx() {
    static Context c;
    return c;
}

int ensure_ready(Context &c) {
    if (c.ready) return WIND_OK;
    WIND_HIP(hipGetDevice(&c.device));
    hipDeviceProp_t prop;
    WIND_HIP(hipGetDeviceProperties(&prop, c.device));
    c.num_cu = prop.multiProcessorCount;
    WIND_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    WIND_HIP(hipEventCreate(&c.ev0));
    WIND_HIP(hipEventCreate(&c.ev1));
    c.ready = true;
    return WIND_OK;
}

int reserve(Pool &p, size_t bytes) {
    if (bytes <= p.bytes) return WIND_OK;
    if (p.ptr) WIND_HIP(hipFree(p.ptr));
    p.ptr = nullptr;
    p.bytes = 0;
    size_t want = std::max(bytes, (size_t)256);
    WIND_HIP(hipMalloc(&p.ptr, want));
    p.bytes = want;
    return WIND_OK;
}

template <class K>
int persistent_grid(const Context &c, K kernel, int block, int nsys, int *grid) {
    int per_cu = 0;
    WIND_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    if (per_cu < 1) per_cu = 1;
    long long want = (long long)c.num_cu * per_cu;
    long long need = ((long long)nsys + block - 1) / block;
    *grid = (int)std::max(1LL, std::min(want, need));
    return WIND_OK;
}

template <class Net, int METHOD>
int launch_small(const Context &c, const wind_network *net, const IntegrateArgs &a, int mode,
                 int kernel, hipStream_t st) {
    typename Net::Params p{};
    if constexpr (sizeof(typename Net::Params) > 1) std::memcpy(&p, net->params.data(), sizeof(p));
    if (kernel == WIND_KERNEL_BLOCK) {
        if (mode != WIND_MODE_EXACT)
            return fail(WIND_ERR_UNSUPPORTED, "the block kernel implements exact mode only");
        int grid = (int)std::min<long long>(a.nsys, (long long)c.num_cu * 32);
        hipLaunchKernelGGL((block_integrate<Net, METHOD>), dim3(grid), dim3(64), 0, st, a, p);
    } else if (mode == WIND_MODE_EXACT) {
        auto k = lane_integrate<Net, METHOD, MathExact>;
        int grid = 0;
        int rc = persistent_grid(c, k, 256, a.nsys, &grid);
        if (rc) return rc;
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    } else {
        auto k = lane_integrate<Net, METHOD, MathFast>;
        int grid = 0;
        int rc = persistent_grid(c, k, 256, a.nsys, &grid);
        if (rc) return rc;
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    }
    WIND_HIP(hipGetLastError());
    return WIND_OK;
}

template <int METHOD>
int dispatch(const Context &c, const wind_network *net, const IntegrateArgs &a, int mode,
             int kernel, hipStream_t st) {
    switch (net->builtin) {
    case WIND_NET_NR_TEST: return launch_small<NetNRTest, METHOD>(c, net, a, mode, kernel, st);
    case WIND_NET_KATZ96: return launch_small<NetKatz96, METHOD>(c, net, a, mode, kernel, st);
    case WIND_NET_HHE6: return launch_small<NetHHe6, METHOD>(c, net, a, mode, kernel, st);
    case WIND_NET_MASS_ACTION: return launch_mass_action<METHOD>(c.num_cu, net, a, mode, kernel, st);
    default: return fail(WIND_ERR_UNSUPPORTED, "unknown network");
    }
}

}  // namespace

int wind_fail(int code, const char *msg) { return fail(code, msg); }

extern "C" {

const char *wind_version(void) { return "wind_amd 0.1 (gfx950)"; }

const char *wind_last_error(void) { return g_last_error.c_str(); }

void wind_options_init(wind_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->method = WIND_SIE;
    o->mode = WIND_MODE_EXACT;
    o->kernel = WIND_KERNEL_AUTO;
}

int wind_network_create_builtin(int builtin, const float *params, int nparams, wind_network **out) {
    if (!out) return fail(WIND_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    wind_network *n = new wind_network();
    n->builtin = builtin;
    switch (builtin) {
    case WIND_NET_NR_TEST:
        n->N = NetNRTest::N;
        n->nconst = NetNRTest::NCONST;
        break;
    case WIND_NET_KATZ96:
        n->N = NetKatz96::N;
        n->nconst = NetKatz96::NCONST;
        break;
    case WIND_NET_HHE6:
        n->N = NetHHe6::N;
        n->nconst = NetHHe6::NCONST;
        if (!params || nparams != 24) {
            delete n;
            return fail(WIND_ERR_ARGUMENT, "hhe6 needs 24 network parameters");
        }
        n->params.assign(params, params + 24);
        break;
    default:
        delete n;
        return fail(WIND_ERR_ARGUMENT, "unknown builtin network (mass-action networks use "
                                       "wind_network_create_mass_action)");
    }
    *out = n;
    return WIND_OK;
}

int wind_network_create_mass_action(int N, int R, const int32_t *rx, const float *k,
                                    wind_network **out) {
    if (!out) return fail(WIND_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (N < 1 || N > kMassActionMaxN) return fail(WIND_ERR_ARGUMENT, "mass-action N out of range");
    if (R < 0 || (R > 0 && (!rx || !k))) return fail(WIND_ERR_ARGUMENT, "bad reaction table");
    for (int i = 0; i < 4 * R; i++)
        if (rx[i] < 0 || rx[i] >= N) return fail(WIND_ERR_ARGUMENT, "species index out of range");
    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    wind_network *n = new wind_network();
    n->builtin = WIND_NET_MASS_ACTION;
    n->N = N;
    n->nconst = 1;
    rc = mass_action_build(n, N, R, rx, k);
    if (rc) {
        wind_network_destroy(n);
        return rc;
    }
    *out = n;
    return WIND_OK;
}

void wind_network_destroy(wind_network *net) {
    if (!net) return;
    mass_action_release(net);
    delete net;
}

int wind_network_size(const wind_network *net) { return net ? net->N : 0; }

int wind_integrate(const wind_network *net, float tnow, float tend, int n_integration_steps,
                   const float *constants, float *equations, int Nsystems, int Neqn,
                   float ABSOLUTE, float RELATIVE, const wind_options *opts_in) {
    g_last_error.clear();
    wind_options opts;
    if (opts_in)
        opts = *opts_in;
    else
        wind_options_init(&opts);
    if (opts.total_steps) *opts.total_steps = 0;
    if (opts.kernel_ms) *opts.kernel_ms = 0.0f;
    if (!net) return fail(WIND_ERR_ARGUMENT, "network is NULL");
    if (opts.method != WIND_SIE && opts.method != WIND_RK2)
        return fail(WIND_ERR_ARGUMENT, "method must be WIND_SIE or WIND_RK2");
    if (opts.mode != WIND_MODE_EXACT && opts.mode != WIND_MODE_FAST)
        return fail(WIND_ERR_ARGUMENT, "mode must be WIND_MODE_EXACT or WIND_MODE_FAST");
    if (opts.controller != WIND_CTRL_REFERENCE && opts.controller != WIND_CTRL_ADAPTIVE)
        return fail(WIND_ERR_ARGUMENT, "controller must be WIND_CTRL_REFERENCE or WIND_CTRL_ADAPTIVE");
    if (Nsystems < 0) return fail(WIND_ERR_ARGUMENT, "Nsystems < 0");
    if (Neqn != net->N)
        return fail(WIND_ERR_ARGUMENT, "Nequations_per_system does not match the network (N = " +
                                           std::to_string(net->N) + ")");
    if (n_integration_steps < 0)
        return fail(WIND_ERR_ARGUMENT, "n_integration_steps < 0 (negative initial step)");
    // As in the reference, systems may share constants when the stride is
    // below the network's constant count (cuda/include/ode_gold.h:1 ships
    // NUM_CONST = 2 for a network that reads 10); the caller's array must
    // hold (Nsystems - 1) * stride + nconst floats.
    const int stride = opts.const_stride > 0 ? opts.const_stride : net->nconst;
    if (Nsystems == 0 || !(tnow < tend)) return WIND_OK;  // no attempt is made (common_gold.c:77)
    if (!constants || !equations) return fail(WIND_ERR_ARGUMENT, "NULL state or constants");

    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    hipStream_t st = opts.stream ? (hipStream_t)opts.stream : c.stream;

    const size_t eq_bytes = (size_t)Nsystems * Neqn * sizeof(float);
    // Reads of constants reach (Nsystems - 1) * stride + nconst floats.
    const size_t c_floats = (size_t)(Nsystems - 1) * stride + net->nconst;
    const size_t c_bytes = c_floats * sizeof(float);

    const float *d_const = constants;
    float *d_eq = equations;
    int32_t *d_steps = opts.steps_out;
    int32_t *d_status = opts.status_out;
    if (!opts.device_pointers) {
        rc = reserve(c.constants, c_bytes);
        if (!rc) rc = reserve(c.equations, eq_bytes);
        if (!rc && opts.steps_out) rc = reserve(c.steps, (size_t)Nsystems * 4);
        if (!rc && opts.status_out) rc = reserve(c.status, (size_t)Nsystems * 4);
        if (rc) return rc;
        WIND_HIP(hipMemcpyAsync(c.constants.ptr, constants, c_bytes, hipMemcpyHostToDevice, st));
        WIND_HIP(hipMemcpyAsync(c.equations.ptr, equations, eq_bytes, hipMemcpyHostToDevice, st));
        d_const = (const float *)c.constants.ptr;
        d_eq = (float *)c.equations.ptr;
        d_steps = opts.steps_out ? (int32_t *)c.steps.ptr : nullptr;
        d_status = opts.status_out ? (int32_t *)c.status.ptr : nullptr;
    }
    rc = reserve(c.counters, 4 * sizeof(unsigned long long));
    if (rc) return rc;
    WIND_HIP(hipMemsetAsync(c.counters.ptr, 0, 4 * sizeof(unsigned long long), st));

    IntegrateArgs a;
    a.tnow = tnow;
    a.tend = tend;
    a.h0 = (tend - tnow) / (float)n_integration_steps;  // sie_gold.c:105
    a.ABS = ABSOLUTE;
    a.REL = RELATIVE;
    a.controller = opts.controller;
    a.constants = d_const;
    a.equations = d_eq;
    a.nsys = Nsystems;
    a.const_stride = stride;
    a.max_attempts = opts.max_attempts > 0 ? opts.max_attempts : kDefaultMaxAttempts;
    a.steps_out = d_steps;
    a.status_out = d_status;
    a.index = nullptr;
    a.counters = (unsigned long long *)c.counters.ptr;

    WIND_HIP(hipEventRecord(c.ev0, st));
    rc = opts.method == WIND_SIE ? dispatch<WIND_SIE>(c, net, a, opts.mode, opts.kernel, st)
                                 : dispatch<WIND_RK2>(c, net, a, opts.mode, opts.kernel, st);
    if (rc) return rc;
    WIND_HIP(hipEventRecord(c.ev1, st));

    unsigned long long total = 0;
    WIND_HIP(hipMemcpyAsync(&total, a.counters + 1, sizeof(total), hipMemcpyDeviceToHost, st));
    if (!opts.device_pointers) {
        WIND_HIP(hipMemcpyAsync(equations, d_eq, eq_bytes, hipMemcpyDeviceToHost, st));
        if (opts.steps_out)
            WIND_HIP(hipMemcpyAsync(opts.steps_out, d_steps, (size_t)Nsystems * 4,
                                    hipMemcpyDeviceToHost, st));
        if (opts.status_out)
            WIND_HIP(hipMemcpyAsync(opts.status_out, d_status, (size_t)Nsystems * 4,
                                    hipMemcpyDeviceToHost, st));
    }
    WIND_HIP(hipStreamSynchronize(st));
    if (opts.kernel_ms) WIND_HIP(hipEventElapsedTime(opts.kernel_ms, c.ev0, c.ev1));
    if (opts.total_steps) *opts.total_steps = (int64_t)total;
    return WIND_OK;
}

}  // extern "C"
