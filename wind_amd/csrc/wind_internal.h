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
    int32_t *r_ptr = nullptr;  // [N+1]  non-zeros by row (for residuals)
    int32_t *r_z = nullptr;    //        indices into j_pos
    // Register-resident dense solver (built in mass_action_build, which
    // documents the formats): the per-thread term streams, word j of thread
    // t at sc_ent[j * 256 + t], sc_len words per thread; the per-tile
    // element masks and offsets into the compact element buffer (12 words
    // per 16x16 tile); sc_cap floats of that buffer.
    int32_t *sc_ent = nullptr;
    uint32_t *sc_tile = nullptr;
    int sc_len = 0, sc_cap = 0;
};

}  // namespace wind

struct wind_plugin_desc;

struct wind_network {
    int builtin = 0;
    int N = 0;
    int nconst = 0;
    std::vector<float> params;
    wind::MassActionDevice ma;
    // Host copies of the mass-action table (kept for introspection).
    std::vector<int32_t> rx;
    std::vector<float> k;
    // Generated networks: the plugin library and its descriptor.
    void *dl = nullptr;
    const wind_plugin_desc *plugin = nullptr;
};

int wind_fail(int code, const char *msg);
int mass_action_build(wind_network *n, int N, int R, const int32_t *rx, const float *k);
void mass_action_release(wind_network *n);
struct SolverChoice {
    int solver = 0;     // wind_solver
    int ir_sweeps = 0;  // refinement sweeps (dense path)
    int block = 0;      // threads per workgroup (dense path); 0 = by N
    int grid = 0;       // workgroups; 0 = fill the device
    float share = 0.0f; // fraction of a full-device grid to use (batches); 0 = all
    int pivot = 0;      // fast mode: partial pivoting in the LU solves (WIND_FLAG_PIVOT)
    int dense_kernel = 0;  // dense path: 0 = by size, 1 = LDS-resident matrix, 2 = register-resident
};

// Workgroups for a persistent kernel: a full device (per_cu per CU), scaled
// by a batch share, capped by an explicit grid and by the system count.
inline int scaled_grid(int num_cu, int per_cu, const SolverChoice &sc, int nsys, int per_block_systems = 1) {
    long long full = (long long)num_cu * (per_cu < 1 ? 1 : per_cu);
    if (sc.share > 0.0f && sc.share < 1.0f) full = (long long)(full * (double)sc.share + 0.999);
    if (sc.grid > 0 && sc.grid < full) full = sc.grid;
    long long need = ((long long)nsys + per_block_systems - 1) / per_block_systems;
    if (full > need) full = need;
    return (int)(full < 1 ? 1 : full);
}
template <int METHOD>
int launch_mass_action(int num_cu, const wind_network *net, const wind::IntegrateArgs &a, int mode,
                       int kernel, const SolverChoice &sc, hipStream_t st);

// Dense LDS/MFMA solver launch (dense.hip); PREC is a kernels_dense.h
// DensePrec value.
template <int METHOD, int PREC>
int launch_dense(int num_cu, const wind::MassActionDevice &d, const wind::IntegrateArgs &a, int ir, size_t lds,
                 int T, const SolverChoice &sc, hipStream_t st);

// Register-resident dense SIE solver (dense.hip): tiles per dimension it is
// built for, whether it takes this network and choice (fast mode, FP32
// tiles, no refinement, two workgroups' LDS per CU), and its launch.
constexpr int kDenseRegMinNT = 2, kDenseRegMaxNT = 10;
// Tiles per dimension at which the register-resident solver is used: at
// three (33 <= N <= 48) it measured 7-10% slower than the LDS solver.
constexpr bool dense_reg_tiles(int NT) { return NT >= kDenseRegMinNT && NT <= kDenseRegMaxNT && NT != 3; }
bool dense_reg_eligible(const wind::MassActionDevice &d, const SolverChoice &sc);
int launch_dense_reg(int num_cu, const wind::MassActionDevice &d, const wind::IntegrateArgs &a, const SolverChoice &sc,
                     hipStream_t st);

// Fast-mode lane kernels of the compiled-in networks (lane_fast.hip,
// compiled with fused multiply-add contraction).
template <class Net, int METHOD>
int launch_lane_fast(int num_cu, const typename Net::Params &p, const wind::IntegrateArgs &a, const SolverChoice &sc,
                     hipStream_t st);

// Start-order hints of a fast lane launch estimated from the initial state
// (lane_fast.hip): hint[nsys]; scratch holds 64 + nsys 32-bit words and is
// cleared by the caller.
template <class Net>
int launch_lane_activity(int num_cu, const typename Net::Params &p, const wind::IntegrateArgs &a, int32_t *hint,
                         unsigned *scratch, hipStream_t st);
// The same from a short probe integration of each system (its first
// attempts, extrapolated to the end of the interval).
template <class Net, int METHOD>
int launch_lane_probe(int num_cu, const typename Net::Params &p, const wind::IntegrateArgs &a, int32_t *hint,
                      unsigned *scratch, hipStream_t st);

// True when launch_mass_action would use the process-wide matrix scratch
// (matrices too large for LDS): only callers holding the context lock.
bool mass_action_needs_scratch(const wind_network *net, int mode);
// Scratch bytes such a launch needs (0 when it uses none), and growing the
// scratch to a given size ahead of launches (batches reserve before they
// enqueue or capture anything).
template <int METHOD>
int mass_action_scratch_bytes(int num_cu, const wind_network *net, int mode, const SolverChoice &sc, int nsys,
                              size_t *bytes);
int mass_action_scratch_reserve(size_t bytes);
// Current address of that scratch (changes when a reserve grows it).
const void *mass_action_scratch_ptr();

// Enqueues one integration kernel on `st` without synchronising (device
// pointers in `a`; a.counters must point at 4 zeroed 64-bit counters).
// Fails with WIND_ERR_UNSUPPORTED for paths that need the shared scratch
// buffer (exact mode on mass-action networks too large for LDS).
int wind_launch_async(const wind_network *net, const wind::IntegrateArgs &a, int method, int mode, int kernel,
                      const SolverChoice &sc, hipStream_t st);
// Device CU count (initialises the library context).
int wind_device_cus(int *num_cu);
// Fills n 64-bit words with zero on `st` (graph- and capture-safe).
void wind_zero_counters(unsigned long long *p, int n, hipStream_t st);
// Longest-first start order of n systems from device cost hints, written to
// index[n] on `st`; count is device scratch of kHintCountWords 32-bit words.
constexpr int kHintCountWords = 128;
// Optional packing of the systems into records in start order (hint_place).
struct PackArgs {
    const float *eq = nullptr, *cst = nullptr;
    int N = 0, NC = 0, stride = 0, R = 0;
    float *rec = nullptr;  // null: no packing
};
void wind_order_by_hint(const int32_t *hint, int n, int32_t *index, unsigned *count, hipStream_t st,
                        const PackArgs *pk = nullptr);
// Whether the kernels launched for (net, mode, kernel) take systems from a
// work queue, so that a start order applies.
bool wind_queue_kernel(const wind_network *net, int mode, int kernel);
