// Host-side internals shared by the translation units of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "wind_types.h"

namespace wind {

constexpr long long kDefaultMaxAttempts = 1LL << 24;
constexpr int kMassActionMaxN = 512;

// Device-side tables of a mass-action network (see kernels_mass_action.h).
struct MassActionDevice {
    int N = 0, R = 0, nnz = 0;
    int32_t *rx = nullptr;     // [4R]
    float *k = nullptr;        // [R]
    int32_t *f_ptr = nullptr;  // [N+1]  CSR over species
    int32_t *f_ent = nullptr;  //        r*2 + (sign < 0)
    int32_t *j_pos = nullptr;  // [nnz]  element index e = col*N + row
    int32_t *j_ptr = nullptr;  // [nnz+1]
    int32_t *j_ent = nullptr;  //        ((r*2 + use_db)*2) + (sign < 0)
};

}  // namespace wind

struct wind_network {
    int builtin = 0;
    int N = 0;
    int nconst = 0;
    std::vector<float> params;
    wind::MassActionDevice ma;
    // Host copies of the mass-action table (kept for introspection).
    std::vector<int32_t> rx;
    std::vector<float> k;
};

int wind_fail(int code, const char *msg);
int mass_action_build(wind_network *n, int N, int R, const int32_t *rx, const float *k);
void mass_action_release(wind_network *n);
template <int METHOD>
int launch_mass_action(int num_cu, const wind_network *net, const wind::IntegrateArgs &a, int mode,
                       int kernel, hipStream_t st);
