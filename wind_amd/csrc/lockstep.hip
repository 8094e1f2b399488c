// Lockstep SIE: the semantics of the reference's host-driven integrator
// (cuda/SIE_host/harness_kernel.cu:248-425 errorLoop, :175-246 solveSystem,
// :21-173 SIE_step).  All systems share one timestep; if any system fails
// the error test every system retries with half the step; after an
// accepted step the step doubles; after 9 consecutive failures the call
// returns with the systems at the last accepted time.  The return value is
// 3 * attempts * Nsystems (harness_kernel.cu:295,424).
//
// Each attempt is one kernel launch over all systems (kernels_lockstep.h)
// that writes the two-half-step result and raises a device flag when a
// system fails; the host reads the 4-byte flag and either commits the
// half-step state or halves the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "kernels_lockstep.h"
#include "plugin_abi.h"
#include "wind_internal.h"

using namespace wind;

namespace {

struct LockstepBuffers {
    std::mutex mu;
    float *C = nullptr, *Y = nullptr, *Y2 = nullptr;
    int *flag = nullptr;
    size_t cb = 0, yb = 0, y2b = 0;
};
LockstepBuffers g_ls;

#define LS_HIP(call)                                                  \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

int grow(float **p, size_t *have, size_t need) {
    if (need <= *have) return WIND_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    LS_HIP(hipMalloc((void **)p, need));
    *have = need;
    return WIND_OK;
}

// One attempt over all systems: launches the network's lockstep kernel.
using AttemptFn = int (*)(const wind_network *net, float tnow, float hh, const float *C, int stride,
                          const float *Y, float *Y2, int nsys, float ABS, float REL, int *flag, int grid);

template <class Net>
int attempt_builtin(const wind_network *net, float tnow, float hh, const float *C, int stride, const float *Y,
                    float *Y2, int nsys, float ABS, float REL, int *flag, int grid) {
    typename Net::Params p{};
    if constexpr (sizeof(typename Net::Params) > 1) std::memcpy(&p, net->params.data(), sizeof(p));
    hipLaunchKernelGGL(lockstep_attempt<Net>, dim3(grid), dim3(256), 0, 0, tnow, hh, C, stride, Y, Y2, nsys, ABS,
                       REL, flag, p);
    LS_HIP(hipGetLastError());
    return WIND_OK;
}

int attempt_plugin(const wind_network *net, float tnow, float hh, const float *C, int stride, const float *Y,
                   float *Y2, int nsys, float ABS, float REL, int *flag, int grid) {
    const int rc = net->plugin->lockstep_attempt(tnow, hh, C, stride, Y, Y2, nsys, ABS, REL, flag, grid);
    return rc ? wind_fail(rc, "generated network: lockstep launch failed") : WIND_OK;
}

int run(const wind_network *net, AttemptFn attempt, float tnow, float tend, int nfix, const float *constants,
        float *equations, int ns, float ABS, float REL, int stride, int64_t *result) {
    const int N = net->N;
    const size_t yb = (size_t)ns * N * sizeof(float);
    const size_t cb = ((size_t)(ns - 1) * stride + net->nconst) * sizeof(float);
    int rc = grow(&g_ls.C, &g_ls.cb, cb);
    if (!rc) rc = grow(&g_ls.Y, &g_ls.yb, yb);
    if (!rc) rc = grow(&g_ls.Y2, &g_ls.y2b, yb);
    if (rc) return rc;
    if (!g_ls.flag) LS_HIP(hipMalloc((void **)&g_ls.flag, sizeof(int)));
    LS_HIP(hipMemcpy(g_ls.C, constants, cb, hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(g_ls.Y, equations, yb, hipMemcpyHostToDevice));
    int dev = 0, ncu = 0;
    LS_HIP(hipGetDevice(&dev));
    LS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = (int)std::max(1LL, std::min((long long)ncu * 8, ((long long)ns + 255) / 256));
    long long nsteps = 0;
    int unsolved = 0;
    float timestep = (tend - tnow) / (float)nfix;
    while (tnow < tend && unsolved < 9) {
        timestep = fminf(timestep, tend - tnow);
        nsteps += 3;
        LS_HIP(hipMemset(g_ls.flag, 0, sizeof(int)));
        rc = attempt(net, tnow, timestep, g_ls.C, stride, g_ls.Y, g_ls.Y2, ns, ABS, REL, g_ls.flag, grid);
        if (rc) return rc;
        int flag = 0;
        LS_HIP(hipMemcpy(&flag, g_ls.flag, sizeof(int), hipMemcpyDeviceToHost));
        if (flag) {
            unsolved++;
            timestep = timestep / 2.0f;
        } else {
            unsolved = 0;
            LS_HIP(hipMemcpy(g_ls.Y, g_ls.Y2, yb, hipMemcpyDeviceToDevice));
            tnow = tnow + timestep;
            timestep = timestep * 2.0f;
        }
    }
    LS_HIP(hipMemcpy(equations, g_ls.Y, yb, hipMemcpyDeviceToHost));
    *result = (int64_t)nsteps * ns;
    return WIND_OK;
}

}  // namespace

extern "C" int wind_integrate_lockstep(const wind_network *net, float tnow, float tend,
                                       int n_integration_steps, const float *constants,
                                       float *equations, int Nsystems, int Neqn, float ABSOLUTE,
                                       float RELATIVE, int const_stride, int64_t *result) {
    if (result) *result = 0;
    if (!net || !result) return wind_fail(WIND_ERR_ARGUMENT, "NULL network or result");
    if (Neqn != net->N) return wind_fail(WIND_ERR_ARGUMENT, "Nequations_per_system does not match");
    if (Nsystems < 0 || n_integration_steps < 0) return wind_fail(WIND_ERR_ARGUMENT, "bad sizes");
    if (Nsystems == 0 || !(tnow < tend)) return WIND_OK;
    if (!constants || !equations) return wind_fail(WIND_ERR_ARGUMENT, "NULL state or constants");
    const int stride = const_stride > 0 ? const_stride : net->nconst;
    std::lock_guard<std::mutex> lock(g_ls.mu);
    AttemptFn fn = nullptr;
    switch (net->builtin) {
    case WIND_NET_NR_TEST: fn = attempt_builtin<NetNRTest>; break;
    case WIND_NET_KATZ96: fn = attempt_builtin<NetKatz96>; break;
    case WIND_NET_HHE6: fn = attempt_builtin<NetHHe6>; break;
    case WIND_NET_PLUGIN: fn = attempt_plugin; break;
    default: return wind_fail(WIND_ERR_UNSUPPORTED, "lockstep SIE is built for the compiled and generated networks");
    }
    return run(net, fn, tnow, tend, n_integration_steps, constants, equations, Nsystems, ABSOLUTE, RELATIVE, stride,
               result);
}
