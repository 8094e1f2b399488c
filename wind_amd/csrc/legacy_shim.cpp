// Drop-in entry points of agurvich/WIND's GPU libraries, compiled once per
// network and integrator like the reference's per-network rebuild
// (python/shell_scripts/precompile_system.sh:7-26, cuda/Makefile:47-60).
// C++ linkage on purpose: WIND's Python frontend looks the functions up by
// their mangled names (python/wind_harness.py:27-40).
//
// WIND loads these libraries from <WIND>/cuda/lib (wind_harness.py:18-54),
// where they are installed as symlinks or copies.  A shim therefore has no
// link-time dependency on the core library: at its first call it finds
// libwind_amd.so (and a generated network's libwind_net.so) from its own
// resolved location, and binds the C-ABI with dlsym.  Search order:
//   1. the file named by the environment variable WIND_AMD_LIBRARY;
//   2. WIND_SHIM_CORE_REL relative to the directory of the shim's real path
//      (symlinks resolved, so an installed symlink finds the build tree);
//   3. WIND_SHIM_CORE_ABS, the core library's absolute path at build time
//      (an installed copy of the shim).
// The plugin of a generated network is looked up the same way, beside the
// shim's real path, then in WIND_SHIM_PLUGIN_ABS.
//
// Build-time definitions:
//   WIND_SHIM_KIND       1 = sie.so, 2 = rk2.so, 3 = sie_host.so
//   WIND_SHIM_NETWORK    a wind_builtin value; 4 = a generated network
//   WIND_SHIM_NUM_CONST  constant stride per system (the reference's NUM_CONST)
//   WIND_SHIM_CORE_REL   core library path relative to the shim's directory
//   WIND_SHIM_CORE_ABS   absolute path of the core library at build time
//   WIND_SHIM_PLUGIN_ABS absolute path of the plugin at build time (network 4)
//   WIND_SHIM_PARAMS     header defining wind_shim_params[] (hhe6) or the
//                        reaction table wind_shim_rx[] / wind_shim_k[] and
//                        WIND_SHIM_N / WIND_SHIM_R (mass action)
#include <dlfcn.h>
#include <limits.h>
#include <stdlib.h>
#include <unistd.h>

#include <cstdio>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wind.h"

#ifdef WIND_SHIM_PARAMS
#include WIND_SHIM_PARAMS
#endif

#ifndef WIND_SHIM_CORE_REL
#define WIND_SHIM_CORE_REL "../libwind_amd.so"
#endif
#ifndef WIND_SHIM_CORE_ABS
#define WIND_SHIM_CORE_ABS ""
#endif
#ifndef WIND_SHIM_PLUGIN_ABS
#define WIND_SHIM_PLUGIN_ABS ""
#endif

namespace {

// The C-ABI of libwind_amd.so, bound at run time.
struct CoreApi {
    void *dl = nullptr;
    int (*create_builtin)(int, const float *, int, wind_network **) = nullptr;
    int (*create_mass_action)(int, int, const int32_t *, const float *, wind_network **) = nullptr;
    int (*create_plugin)(const char *, wind_network **) = nullptr;
    void (*options_init)(wind_options *) = nullptr;
    int (*integrate)(const wind_network *, float, float, int, const float *, float *, int, int, float, float,
                     const wind_options *) = nullptr;
    int (*integrate_lockstep)(const wind_network *, float, float, int, const float *, float *, int, int, float,
                              float, int, int64_t *) = nullptr;
    const char *(*last_error)(void) = nullptr;
};

// Directory of this shim's real file (symlinks resolved).
std::string shim_dir() {
    Dl_info info;
    if (!dladdr((void *)&shim_dir, &info) || !info.dli_fname) return ".";
    char buf[PATH_MAX];
    const char *real = realpath(info.dli_fname, buf);
    const std::string self(real ? real : info.dli_fname);
    const size_t slash = self.rfind('/');
    return slash == std::string::npos ? std::string(".") : self.substr(0, slash);
}

bool readable(const std::string &p) { return !p.empty() && access(p.c_str(), R_OK) == 0; }

// First readable candidate, or "" (the caller reports every candidate).
std::string first_readable(const std::vector<std::string> &cand) {
    for (const std::string &p : cand)
        if (readable(p)) return p;
    return "";
}

std::string joined(const std::vector<std::string> &cand) {
    std::string s;
    for (const std::string &p : cand)
        if (!p.empty()) s += (s.empty() ? "" : ", ") + p;
    return s;
}

template <class F>
bool bind(void *dl, const char *name, F &fn) {
    fn = reinterpret_cast<F>(dlsym(dl, name));
    if (!fn) std::fprintf(stderr, "wind_amd: %s not found in the core library\n", name);
    return fn != nullptr;
}

const CoreApi *core() {
    static std::once_flag once;
    static CoreApi api;
    static bool ok = false;
    std::call_once(once, [] {
        const char *env = getenv("WIND_AMD_LIBRARY");
        const std::vector<std::string> cand = {env ? std::string(env) : std::string(),
                                               shim_dir() + "/" + WIND_SHIM_CORE_REL, WIND_SHIM_CORE_ABS};
        const std::string path = first_readable(cand);
        if (path.empty()) {
            std::fprintf(stderr, "wind_amd: libwind_amd.so not found (tried %s)\n", joined(cand).c_str());
            return;
        }
        api.dl = dlopen(path.c_str(), RTLD_NOW | RTLD_GLOBAL);
        if (!api.dl) {
            std::fprintf(stderr, "wind_amd: cannot load %s: %s\n", path.c_str(), dlerror());
            return;
        }
        ok = bind(api.dl, "wind_network_create_builtin", api.create_builtin) &&
             bind(api.dl, "wind_network_create_mass_action", api.create_mass_action) &&
             bind(api.dl, "wind_network_create_plugin", api.create_plugin) &&
             bind(api.dl, "wind_options_init", api.options_init) && bind(api.dl, "wind_integrate", api.integrate) &&
             bind(api.dl, "wind_integrate_lockstep", api.integrate_lockstep) &&
             bind(api.dl, "wind_last_error", api.last_error);
    });
    return ok ? &api : nullptr;
}

wind_network *shim_network() {
    static std::once_flag once;
    static wind_network *net = nullptr;
    std::call_once(once, [] {
        const CoreApi *A = core();
        if (!A) return;
        int rc;
#if WIND_SHIM_NETWORK == 4
        const std::vector<std::string> cand = {shim_dir() + "/libwind_net.so", WIND_SHIM_PLUGIN_ABS};
        const std::string path = first_readable(cand);
        if (path.empty()) {
            std::fprintf(stderr, "wind_amd: libwind_net.so not found (tried %s)\n", joined(cand).c_str());
            return;
        }
        rc = A->create_plugin(path.c_str(), &net);
#elif WIND_SHIM_NETWORK == 3
        rc = A->create_mass_action(WIND_SHIM_N, WIND_SHIM_R, wind_shim_rx, wind_shim_k, &net);
#elif WIND_SHIM_NETWORK == 2
        rc = A->create_builtin(WIND_SHIM_NETWORK, wind_shim_params, 24, &net);
#else
        rc = A->create_builtin(WIND_SHIM_NETWORK, nullptr, 0, &net);
#endif
        if (rc) {
            std::fprintf(stderr, "wind_amd: network setup failed: %s\n", A->last_error());
            net = nullptr;
        }
    });
    return net;
}

// WIND's output loop (python/ode_systems/ode_base.py:315-397) passes the same
// state array for every output step.  For a batch larger than the device
// runs at once (2 x 256 x 256 lanes on the fast kernel; below that every
// system starts immediately and order cannot matter), a call on the array
// of the call before starts its systems longest first by that call's
// per-system sub-step counts (wind_options.cost_hint).  Results do not
// depend on the order.
constexpr int kOrderMinSystems = 1 << 17;

struct LastCall {
    std::mutex mu;
    const float *equations = nullptr;
    int nsys = 0;
    bool valid = false;
    std::vector<int32_t> prev, cur;
};

LastCall &last_call() {
    static LastCall l;
    return l;
}

int shim_integrate(int method, float tnow, float tend, int n_integration_steps, float *constants,
                   float *equations, int Nsystems, int Nequations_per_system, float ABSOLUTE,
                   float RELATIVE) {
    wind_network *net = shim_network();
    if (!net) return -1;
    const CoreApi *A = core();
    wind_options o;
    A->options_init(&o);
    o.method = method;
    o.mode = WIND_MODE_EXACT;
    o.const_stride = WIND_SHIM_NUM_CONST;
    int64_t total = 0;
    o.total_steps = &total;
    LastCall &l = last_call();
    std::lock_guard<std::mutex> lock(l.mu);
    const bool track = Nsystems >= kOrderMinSystems;
    if (track) {
        if (l.valid && l.equations == equations && l.nsys == Nsystems) o.cost_hint = l.prev.data();
        l.cur.resize((size_t)Nsystems);
        o.steps_out = l.cur.data();
    }
    int rc = A->integrate(net, tnow, tend, n_integration_steps, constants, equations, Nsystems,
                            Nequations_per_system, ABSOLUTE, RELATIVE, &o);
    if (track) {
        l.valid = rc == 0;
        l.equations = equations;
        l.nsys = Nsystems;
        std::swap(l.prev, l.cur);
    }
    if (rc) {
        std::fprintf(stderr, "wind_amd: %s\n", A->last_error());
        return -1;
    }
    // The reference accumulates into an int (cuda/common/kernel.cu:307-309).
    return (int)(unsigned int)(uint64_t)total;
}

}  // namespace

#if WIND_SHIM_KIND == 1 || WIND_SHIM_KIND == 2
// cuda/common/harness.cu:5-106
int cudaIntegrateSystem(float tnow, float tend, int n_integration_steps, float *constants,
                        float *equations, int Nsystems, int Nequations_per_system, float ABSOLUTE,
                        float RELATIVE) {
    return shim_integrate(WIND_SHIM_KIND == 1 ? WIND_SIE : WIND_RK2, tnow, tend, n_integration_steps,
                          constants, equations, Nsystems, Nequations_per_system, ABSOLUTE, RELATIVE);
}
#endif

#if WIND_SHIM_KIND == 3
// cuda/SIE_host/harness_kernel.cu:427-532 (lockstep SIE)
int cudaIntegrateSIE(float tnow, float tend, int n_integration_steps, float *constants,
                     float *equations, int Nsystems, int Nequations_per_system, float ABSOLUTE,
                     float RELATIVE) {
    wind_network *net = shim_network();
    if (!net) return -1;
    const CoreApi *A = core();
    int64_t result = 0;
    int rc = A->integrate_lockstep(net, tnow, tend, n_integration_steps, constants, equations,
                                     Nsystems, Nequations_per_system, ABSOLUTE, RELATIVE,
                                     WIND_SHIM_NUM_CONST, &result);
    if (rc) {
        std::fprintf(stderr, "wind_amd: %s\n", A->last_error());
        return -1;
    }
    return (int)(unsigned int)(uint64_t)result;
}

// The reference initialises cuBLAS here (python/wind_harness.py:27,139);
// this library needs no BLAS handle, so it only sets up the network.
void initializeCublasExternally() { (void)shim_network(); }
#endif
