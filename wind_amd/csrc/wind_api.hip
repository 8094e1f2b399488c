// Host side of the wind_amd C-ABI (include/wind.h): device context and
// buffer pool, network handles, argument checks, kernel dispatch.
//
// Replaces the per-call host harness of the reference
// (cuda/common/harness.cu:5-106), which allocates, copies, launches, copies
// back and frees on every call.  Here device buffers are pooled per process
// (grow-only, guarded by a mutex) and reused across calls, and callers that
// keep their state on the device pass device pointers and skip the copies.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "kernels_small.h"
#include "plugin_abi.h"
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

// Autotuned kernel shape per (network, method, mode, controller, solver,
// refinement sweeps), with its measured cost per system.  Guarded by the
// context lock.
struct TuneKey {
    const wind_network *net;
    int method, mode, controller, solver, ir, pivot;
    bool operator<(const TuneKey &o) const {
        return std::tie(net, method, mode, controller, solver, ir, pivot) <
               std::tie(o.net, o.method, o.mode, o.controller, o.solver, o.ir, o.pivot);
    }
};
struct TuneVal {
    SolverChoice sc;
    double sec_per_sys = 0.0;
};
std::map<TuneKey, TuneVal> g_tuned;

// What one wind_integrate call uses on the device: its stream, timing
// events and pooled buffers (grow-only).  Calls take a free slot, so that
// callers on several threads run concurrently.
struct CallSlot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Pool constants, equations, steps, status, counters;
    Pool hint, order;  // cost hints copied from the host; start order and class counts
    Pool records;      // systems packed in start order (lane kernels of compiled-in networks)
    Pool activity;     // initial-state work estimates (start order without caller hints)
    Pool trace;        // per-attempt trace arrays of host-pointer calls
    bool busy = false;
};

struct Context {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<CallSlot *> slots;  // wind_integrate call slots (never freed)
    Pool constants, counters;       // batches and the output-step driver
    Pool hint, order;
    Pool out_steps;    // output-step driver: per-system sub-steps of the last interval
    // batch execution
    std::vector<hipStream_t> bucket_streams;
    std::vector<hipEvent_t> bucket_events;
    hipEvent_t fork = nullptr;
    Pool tune_state, batch_counters;
    std::vector<Pool> bucket_order;  // per bucket: start order from its cost hints
    // output-step driver
    Pool snapshots, out_counters;
    std::vector<hipEvent_t> out_events;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
};

// Longest-first start order from per-system cost hints: a counting sort
// into classes of four per octave of the hint, larger classes first (a
// non-positive hint is class 0, last).  Order within a class is arbitrary;
// the kernels with a work queue take systems in the resulting order.
constexpr int kHintClasses = 128;
constexpr int kHintBlocks = 1024;

__device__ inline int hint_class(int32_t h) {
    if (h <= 0) return 0;
    const unsigned u = (unsigned)h;
    const int lg = 31 - __clz(u);
    const int sub = lg >= 2 ? (int)((u >> (lg - 2)) & 3u) : (int)((u << (2 - lg)) & 3u);
    return 1 + 4 * lg + sub;
}

// Each block counts the classes of its contiguous chunk of systems.
__global__ void hint_count(const int32_t *hint, int n, unsigned *count) {
    __shared__ unsigned h[kHintClasses];
    for (int i = threadIdx.x; i < kHintClasses; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&h[hint_class(hint[i])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kHintClasses; i += blockDim.x)
        if (h[i]) atomicAdd(&count[i], h[i]);
}

// Class counts to start positions, the largest class first (one thread).
__global__ void hint_starts(unsigned *count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned run = 0;
    for (int c = kHintClasses - 1; c >= 0; c--) {
        const unsigned k = count[c];
        count[c] = run;
        run += k;
    }
}

// Each block reserves its chunk's range in every class, then places its
// systems there; with pk.rec, it also packs each system's state, constants
// and id into its record at that position (read in index order, one record
// written per system).
__global__ void hint_place(const int32_t *hint, int n, unsigned *start, int32_t *index, PackArgs pk) {
    __shared__ unsigned h[kHintClasses], base[kHintClasses];
    for (int i = threadIdx.x; i < kHintClasses; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&h[hint_class(hint[i])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kHintClasses; i += blockDim.x) {
        base[i] = h[i] ? atomicAdd(&start[i], h[i]) : 0u;
        h[i] = 0;
    }
    __syncthreads();
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int c = hint_class(hint[i]);
        const unsigned pos = base[c] + atomicAdd(&h[c], 1u);
        index[pos] = i;
        if (pk.rec) {
            // whole 16-byte stores, so that a record is written as full sectors
            float4 *r = reinterpret_cast<float4 *>(pk.rec + (size_t)pos * pk.R);
            auto word = [&](int j) {
                return j < pk.N ? pk.eq[(size_t)i * pk.N + j]
                       : j < pk.N + pk.NC ? pk.cst[(size_t)i * pk.stride + (j - pk.N)]
                       : j == pk.N + pk.NC ? __int_as_float(i) : 0.0f;
            };
            for (int v = 0; v < pk.R / 4; v++) r[v] = make_float4(word(4 * v), word(4 * v + 1), word(4 * v + 2), word(4 * v + 3));
        }
    }
}

__global__ void zero_u64(unsigned long long *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

Context &ctx() {
    static Context c;
    return c;
}

// Launches whose matrices exceed LDS share the process-wide matrix scratch
// (mass_action_scratch_reserve): a call that uses it holds this lock from
// the reserve through its stream synchronisation.  Lock order: the context
// lock first.
std::mutex g_scratch_mu;

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

// A free call slot (created on first need); caller holds the context lock.
int acquire_slot(Context &c, CallSlot **out) {
    for (CallSlot *sl : c.slots)
        if (!sl->busy) {
            sl->busy = true;
            *out = sl;
            return WIND_OK;
        }
    CallSlot *sl = new CallSlot();
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&sl->ev0) != hipSuccess || hipEventCreate(&sl->ev1) != hipSuccess) {
        delete sl;
        return fail(WIND_ERR_DEVICE, "cannot create a stream or events for a call slot");
    }
    sl->busy = true;
    c.slots.push_back(sl);
    *out = sl;
    return WIND_OK;
}

struct SlotGuard {
    Context &c;
    CallSlot *sl;
    ~SlotGuard() {
        std::lock_guard<std::mutex> lock(c.mu);
        sl->busy = false;
    }
};

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

// Writes the longest-first start order of `n` systems from device hints
// into pool `order` (the permutation, then the class counts) on stream st.
int order_by_hint(Pool &order, const int32_t *d_hint, int n, hipStream_t st, const int32_t **index_out,
                  const PackArgs *pk = nullptr) {
    const size_t index_bytes = ((size_t)n * 4 + 255) & ~(size_t)255;
    const int rc = reserve(order, index_bytes + kHintClasses * sizeof(unsigned));
    if (rc) return rc;
    int32_t *index = (int32_t *)order.ptr;
    wind_order_by_hint(d_hint, n, index, (unsigned *)((char *)order.ptr + index_bytes), st, pk);
    WIND_HIP(hipGetLastError());
    *index_out = index;
    return WIND_OK;
}

bool queue_kernel(const wind_network *net, const wind_options &o) { return wind_queue_kernel(net, o.mode, o.kernel); }

// Batches at least this large start in the order of an initial-state work
// estimate when the caller gives no hints (fast mode, compiled-in
// networks): the persistent fast lane grid holds 2 x 256 x 256 systems at
// once, and smaller batches start every system immediately.
constexpr int kAutoOrderMin = 1 << 17;

// Start-order estimate without caller hints: a short probe integration of
// every system (true) or the initial state's fastest rate of change (false).
constexpr bool kAutoOrderProbe = true;

template <class Net>
int activity_hints_for(int num_cu, const wind_network *net, const IntegrateArgs &a, int method, int32_t *hint,
                       unsigned *scratch, hipStream_t st) {
    typename Net::Params p{};
    if constexpr (sizeof(typename Net::Params) > 1) std::memcpy(&p, net->params.data(), sizeof(p));
    if (!kAutoOrderProbe) return launch_lane_activity<Net>(num_cu, p, a, hint, scratch, st);
    return method == WIND_SIE ? launch_lane_probe<Net, WIND_SIE>(num_cu, p, a, hint, scratch, st)
                              : launch_lane_probe<Net, WIND_RK2>(num_cu, p, a, hint, scratch, st);
}

int activity_hints(int num_cu, const wind_network *net, const IntegrateArgs &a, int method, int32_t *hint,
                   unsigned *scratch, hipStream_t st) {
    switch (net->builtin) {
    case WIND_NET_NR_TEST: return activity_hints_for<NetNRTest>(num_cu, net, a, method, hint, scratch, st);
    case WIND_NET_KATZ96: return activity_hints_for<NetKatz96>(num_cu, net, a, method, hint, scratch, st);
    case WIND_NET_HHE6: return activity_hints_for<NetHHe6>(num_cu, net, a, method, hint, scratch, st);
    default: return fail(WIND_ERR_UNSUPPORTED, "no work estimate for this network");
    }
}

template <class K>
int persistent_grid(int num_cu, K kernel, int block, int nsys, int *grid, const SolverChoice &sc) {
    int per_cu = 0;
    WIND_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    *grid = scaled_grid(num_cu, per_cu, sc, nsys, block);
    return WIND_OK;
}

template <class Net, int METHOD>
int launch_small(int num_cu, const wind_network *net, const IntegrateArgs &a, int mode,
                 int kernel, const SolverChoice &sc, hipStream_t st) {
    typename Net::Params p{};
    if constexpr (sizeof(typename Net::Params) > 1) std::memcpy(&p, net->params.data(), sizeof(p));
    if (kernel == WIND_KERNEL_BLOCK) {
        if (mode != WIND_MODE_EXACT)
            return fail(WIND_ERR_UNSUPPORTED, "the block kernel implements exact mode only");
        if (a.trace.capacity > 0) return fail(WIND_ERR_UNSUPPORTED, "trace: the block kernel has no trace");
        const int grid = scaled_grid(num_cu, 32, sc, a.nsys);
        hipLaunchKernelGGL((block_integrate<Net, METHOD>), dim3(grid), dim3(64), 0, st, a, p);
    } else if (mode == WIND_MODE_EXACT) {
        auto k = a.trace.capacity > 0 ? lane_integrate<Net, METHOD, MathExact, true>
                                      : lane_integrate<Net, METHOD, MathExact>;
        int grid = 0;
        int rc = persistent_grid(num_cu, k, 256, a.nsys, &grid, sc);
        if (rc) return rc;
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, a, p);
    } else {
        return launch_lane_fast<Net, METHOD>(num_cu, p, a, sc, st);
    }
    WIND_HIP(hipGetLastError());
    return WIND_OK;
}

template <int METHOD>
int dispatch(int num_cu, const wind_network *net, const IntegrateArgs &a_in, int mode,
             int kernel, const SolverChoice &sc, hipStream_t st) {
    IntegrateArgs a = a_in;
    a.pivot = sc.pivot;
    switch (net->builtin) {
    case WIND_NET_NR_TEST: return launch_small<NetNRTest, METHOD>(num_cu, net, a, mode, kernel, sc, st);
    case WIND_NET_KATZ96: return launch_small<NetKatz96, METHOD>(num_cu, net, a, mode, kernel, sc, st);
    case WIND_NET_HHE6: return launch_small<NetHHe6, METHOD>(num_cu, net, a, mode, kernel, sc, st);
    case WIND_NET_MASS_ACTION: return launch_mass_action<METHOD>(num_cu, net, a, mode, kernel, sc, st);
    case WIND_NET_PLUGIN: {
        if (a.trace.capacity > 0) return fail(WIND_ERR_UNSUPPORTED, "trace: generated networks have no trace");
        // Matrices beyond LDS live in the process-wide scratch (callers hold
        // the context lock; wind_launch_async and the allocation-free entry
        // point refuse such launches).
        size_t need = 0;
        int rc = net->plugin->scratch_bytes(a.nsys, METHOD, mode, kernel, num_cu, sc.share, sc.grid, &need);
        if (rc) return fail(rc, "generated network: no launch shape");
        float *scratch = nullptr;
        if (need) {
            rc = mass_action_scratch_reserve(need);
            if (rc) return rc;
            scratch = (float *)mass_action_scratch_ptr();
        }
        rc = net->plugin->launch(&a, METHOD, mode, kernel, num_cu, sc.share, sc.grid, scratch, need, st);
        if (rc == WIND_ERR_UNSUPPORTED)
            return fail(rc, "generated network: no kernel for this mode and kernel family at N = " +
                                std::to_string(net->N) + " (lane kernels: N <= 16; 64-thread block kernel: exact "
                                "mode, N <= 64)");
        return rc ? fail(rc, "generated network: kernel launch failed") : WIND_OK;
    }
    default: return fail(WIND_ERR_UNSUPPORTED, "unknown network");
    }
}

// Device scratch bytes a launch needs (matrices beyond LDS), 0 if none.
template <int METHOD>
int launch_scratch_bytes(int num_cu, const wind_network *net, int mode, int kernel, const SolverChoice &sc, int nsys,
                         size_t *bytes) {
    *bytes = 0;
    if (net->builtin == WIND_NET_MASS_ACTION) return mass_action_scratch_bytes<METHOD>(num_cu, net, mode, sc, nsys, bytes);
    if (net->builtin == WIND_NET_PLUGIN) {
        const int rc = net->plugin->scratch_bytes(nsys, METHOD, mode, kernel, num_cu, sc.share, sc.grid, bytes);
        return rc ? fail(rc, "generated network: no launch shape") : WIND_OK;
    }
    return WIND_OK;
}

int scratch_bytes_for(int num_cu, const wind_network *net, int method, int mode, int kernel, const SolverChoice &sc,
                      int nsys, size_t *bytes) {
    return method == WIND_SIE ? launch_scratch_bytes<WIND_SIE>(num_cu, net, mode, kernel, sc, nsys, bytes)
                              : launch_scratch_bytes<WIND_RK2>(num_cu, net, mode, kernel, sc, nsys, bytes);
}

}  // namespace

int wind_fail(int code, const char *msg) { return fail(code, msg); }

// Argument checks shared by wind_integrate and wind_integrate_noalloc; also
// clears the optional outputs.
static int check_integrate_args(const wind_network *net, const wind_options &opts, int Nsystems, int Neqn,
                                int n_integration_steps) {
    if (opts.total_steps) *opts.total_steps = 0;
    if (opts.kernel_ms) *opts.kernel_ms = 0.0f;
    if (!net) return fail(WIND_ERR_ARGUMENT, "network is NULL");
    if (opts.method != WIND_SIE && opts.method != WIND_RK2)
        return fail(WIND_ERR_ARGUMENT, "method must be WIND_SIE or WIND_RK2");
    if (opts.mode != WIND_MODE_EXACT && opts.mode != WIND_MODE_FAST)
        return fail(WIND_ERR_ARGUMENT, "mode must be WIND_MODE_EXACT or WIND_MODE_FAST");
    if (opts.controller != WIND_CTRL_REFERENCE && opts.controller != WIND_CTRL_ADAPTIVE)
        return fail(WIND_ERR_ARGUMENT, "controller must be WIND_CTRL_REFERENCE or WIND_CTRL_ADAPTIVE");
    if (opts.solver < WIND_SOLVER_AUTO || opts.solver > WIND_SOLVER_MFMA_BF16X3)
        return fail(WIND_ERR_ARGUMENT, "unknown solver");
    if (opts.ir_sweeps < 0 || opts.ir_sweeps > 4) return fail(WIND_ERR_ARGUMENT, "ir_sweeps must be 0-4");
    if ((opts.flags & WIND_FLAG_PIVOT) && opts.mode != WIND_MODE_FAST)
        return fail(WIND_ERR_ARGUMENT, "WIND_FLAG_PIVOT applies to fast mode (exact mode is the reference's "
                                       "unpivoted Gauss-Jordan)");
    if (Nsystems < 0) return fail(WIND_ERR_ARGUMENT, "Nsystems < 0");
    if (Neqn != net->N)
        return fail(WIND_ERR_ARGUMENT, "Nequations_per_system does not match the network (N = " +
                                           std::to_string(net->N) + ")");
    if (n_integration_steps < 0)
        return fail(WIND_ERR_ARGUMENT, "n_integration_steps < 0 (negative initial step)");
    return WIND_OK;
}

// Device-side per-attempt trace of a wind_integrate call: checks the
// wind_trace, and for host pointers stages it through pool `pool` (replay
// arrays copied in on `st`).  Outputs to copy back afterwards are listed in
// *back (device source, host destination, bytes).
struct TraceCopy {
    const void *src;
    void *dst;
    size_t bytes;
};
static int setup_trace(Pool &pool, const wind_options &o, int n, int N, TraceArgs *tr, std::vector<TraceCopy> *back,
                       hipStream_t st) {
    const wind_trace &t = *o.trace;
    if (t.capacity <= 0) return fail(WIND_ERR_ARGUMENT, "trace: capacity must be positive");
    const bool replay = t.h || t.accept || t.count;
    if (replay && !(t.h && t.accept && t.count)) return fail(WIND_ERR_ARGUMENT, "trace: replay needs h, accept and count");
    if (!replay && !t.y1 && !t.y2 && !t.flag) return fail(WIND_ERR_ARGUMENT, "trace: nothing to replay or record");
    const size_t slots = (size_t)n * t.capacity;
    tr->capacity = t.capacity;
    if (t.y_start && !replay) return fail(WIND_ERR_ARGUMENT, "trace: y_start needs a replay (h, accept, count)");
    if (o.device_pointers) {
        tr->h = t.h;
        tr->accept = t.accept;
        tr->count = t.count;
        tr->y_start = t.y_start;
        tr->y1 = t.y1;
        tr->y2 = t.y2;
        tr->flag = t.flag;
        return WIND_OK;
    }
    if (replay)
        for (int s = 0; s < n; s++)
            if (t.count[s] < 0 || t.count[s] > t.capacity)
                return fail(WIND_ERR_ARGUMENT, "trace: count[" + std::to_string(s) + "] outside [0, capacity]");
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bh = replay ? up(slots * 4) : 0, bc = replay ? up((size_t)n * 4) : 0;
    const size_t by = up(slots * N * 4), bf = up(slots * 4);
    const size_t total = 2 * bh + bc + (t.y_start ? by : 0) + (t.y1 ? by : 0) + (t.y2 ? by : 0) + (t.flag ? bf : 0);
    int rc = reserve(pool, total);
    if (rc) return rc;
    char *p = (char *)pool.ptr;
    if (replay) {
        WIND_HIP(hipMemcpyAsync(p, t.h, slots * 4, hipMemcpyHostToDevice, st));
        tr->h = (const float *)p;
        p += bh;
        WIND_HIP(hipMemcpyAsync(p, t.accept, slots * 4, hipMemcpyHostToDevice, st));
        tr->accept = (const int32_t *)p;
        p += bh;
        WIND_HIP(hipMemcpyAsync(p, t.count, (size_t)n * 4, hipMemcpyHostToDevice, st));
        tr->count = (const int32_t *)p;
        p += bc;
        if (t.y_start) {
            WIND_HIP(hipMemcpyAsync(p, t.y_start, slots * N * 4, hipMemcpyHostToDevice, st));
            tr->y_start = (const float *)p;
            p += by;
        }
    }
    if (t.y1) {
        tr->y1 = (float *)p;
        back->push_back({p, t.y1, slots * N * 4});
        p += by;
    }
    if (t.y2) {
        tr->y2 = (float *)p;
        back->push_back({p, t.y2, slots * N * 4});
        p += by;
    }
    if (t.flag) {
        tr->flag = (int32_t *)p;
        back->push_back({p, t.flag, slots * 4});
    }
    return WIND_OK;
}

// Whether a launch of this size starts in the order of an initial-state
// work estimate (no caller hints; fast lane kernels of compiled-in networks).
static bool auto_order(const wind_network *net, const wind_options &opts, int Nsystems);

static inline uint32_t __float_as_uint_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

extern "C" {

const char *wind_version(void) { return "wind_amd 0.2 (gfx950)"; }

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

int wind_network_create_plugin(const char *path, wind_network **out) {
    if (!out) return fail(WIND_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!path) return fail(WIND_ERR_ARGUMENT, "path is NULL");
    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    void *dl = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!dl) return fail(WIND_ERR_ARGUMENT, (std::string("cannot load network plugin: ") + dlerror()).c_str());
    auto get = (wind_plugin_get_fn)dlsym(dl, "wind_plugin_get");
    const wind_plugin_desc *d = get ? get() : nullptr;
    if (!d || d->abi != WIND_PLUGIN_ABI || d->N < 1 || d->nconst < 0 || !d->launch || !d->lockstep_attempt) {
        dlclose(dl);
        return fail(WIND_ERR_ARGUMENT, "not a network plugin of this library version");
    }
    wind_network *n = new wind_network();
    n->builtin = WIND_NET_PLUGIN;
    n->N = d->N;
    n->nconst = d->nconst;
    n->dl = dl;
    n->plugin = d;
    *out = n;
    return WIND_OK;
}

void wind_network_destroy(wind_network *net) {
    if (!net) return;
    {
        // Cached batch state must not outlive the network: the captured graph
        // holds its device tables, and a later network may reuse the address.
        Context &c = ctx();
        std::lock_guard<std::mutex> lock(c.mu);
        if (c.graph_exec && std::find(c.graph_key.begin(), c.graph_key.end(), (uint64_t)(uintptr_t)net) !=
                                c.graph_key.end()) {
            (void)hipDeviceSynchronize();
            (void)hipGraphExecDestroy(c.graph_exec);
            c.graph_exec = nullptr;
            c.graph_key.clear();
        }
        for (auto it = g_tuned.begin(); it != g_tuned.end();)
            it = it->first.net == net ? g_tuned.erase(it) : std::next(it);
    }
    mass_action_release(net);
    if (net->dl) dlclose(net->dl);
    delete net;
}

int wind_network_size(const wind_network *net) { return net ? net->N : 0; }

int wind_network_nconst(const wind_network *net) { return net ? net->nconst : 0; }

int wind_integrate_outputs(const wind_network *net, double tnow, double tend, int n_output_steps,
                           int n_integration_steps, const float *constants, float *equations, int Nsystems,
                           int Neqn, float ABSOLUTE, float RELATIVE, const wind_options *opts_in,
                           float *equations_over_time, double *times, int64_t *nsteps, double *step_ms) {
    g_last_error.clear();
    wind_options opts;
    if (opts_in)
        opts = *opts_in;
    else
        wind_options_init(&opts);
    if (!net) return fail(WIND_ERR_ARGUMENT, "network is NULL");
    if (opts.method != WIND_SIE && opts.method != WIND_RK2) return fail(WIND_ERR_ARGUMENT, "bad method");
    if (opts.mode != WIND_MODE_EXACT && opts.mode != WIND_MODE_FAST) return fail(WIND_ERR_ARGUMENT, "bad mode");
    if (opts.controller != WIND_CTRL_REFERENCE && opts.controller != WIND_CTRL_ADAPTIVE)
        return fail(WIND_ERR_ARGUMENT, "bad controller");
    if (opts.solver < WIND_SOLVER_AUTO || opts.solver > WIND_SOLVER_MFMA_BF16X3)
        return fail(WIND_ERR_ARGUMENT, "unknown solver");
    if (opts.ir_sweeps < 0 || opts.ir_sweeps > 4) return fail(WIND_ERR_ARGUMENT, "ir_sweeps must be 0-4");
    if (n_output_steps < 1) return fail(WIND_ERR_ARGUMENT, "n_output_steps < 1");
    if (opts.trace) return fail(WIND_ERR_ARGUMENT, "trace: supported by wind_integrate only");
    if ((opts.flags & WIND_FLAG_PIVOT) && opts.mode != WIND_MODE_FAST)
        return fail(WIND_ERR_ARGUMENT, "WIND_FLAG_PIVOT applies to fast mode");
    if (n_integration_steps < 0) return fail(WIND_ERR_ARGUMENT, "n_integration_steps < 0");
    if (Nsystems < 0) return fail(WIND_ERR_ARGUMENT, "Nsystems < 0");
    if (Neqn != net->N) return fail(WIND_ERR_ARGUMENT, "Nequations_per_system does not match the network");
    if (Nsystems > 0 && (!constants || !equations)) return fail(WIND_ERR_ARGUMENT, "NULL state or constants");
    const int stride = opts.const_stride > 0 ? opts.const_stride : net->nconst;
    // Output times as the reference's loop forms them (ode_base.py:324-383):
    // dt = (tend - tnow) / n_output_steps, tcur += dt, in double.
    const double dt = (tend - tnow) / n_output_steps;
    {
        double tc = tnow;
        if (times) times[0] = tc;
        for (int i = 0; i < n_output_steps; i++) {
            tc += dt;
            if (times) times[i + 1] = tc;
        }
    }
    if (nsteps)
        for (int i = 0; i < n_output_steps; i++) nsteps[i] = 0;
    if (step_ms)
        for (int i = 0; i < n_output_steps; i++) step_ms[i] = 0.0;
    const size_t eq_floats = (size_t)Nsystems * Neqn;
    const size_t eq_bytes = eq_floats * sizeof(float);
    if (Nsystems == 0) return WIND_OK;

    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    hipStream_t st = opts.stream ? (hipStream_t)opts.stream : c.stream;
    const size_t c_bytes = ((size_t)(Nsystems - 1) * stride + net->nconst) * sizeof(float);
    // Snapshot k (k = 0 .. n_output_steps) of the state lives in slot k of a
    // device buffer; step k copies slot k - 1 into slot k and integrates it in
    // place, so the state never leaves the device between output steps.
    rc = reserve(c.snapshots, eq_bytes * (size_t)(n_output_steps + 1));
    if (!rc) rc = reserve(c.out_counters, (size_t)n_output_steps * 4 * sizeof(unsigned long long));
    if (!rc && !opts.device_pointers) rc = reserve(c.constants, c_bytes);
    if (rc) return rc;
    while ((int)c.out_events.size() < n_output_steps + 1) {
        hipEvent_t e;
        WIND_HIP(hipEventCreate(&e));
        c.out_events.push_back(e);
    }
    float *snap = (float *)c.snapshots.ptr;
    const float *d_const = constants;
    if (!opts.device_pointers) {
        WIND_HIP(hipMemcpyAsync(c.constants.ptr, constants, c_bytes, hipMemcpyHostToDevice, st));
        WIND_HIP(hipMemcpyAsync(snap, equations, eq_bytes, hipMemcpyHostToDevice, st));
        d_const = (const float *)c.constants.ptr;
    } else {
        WIND_HIP(hipMemcpyAsync(snap, equations, eq_bytes, hipMemcpyDeviceToDevice, st));
    }
    unsigned long long *cnt = (unsigned long long *)c.out_counters.ptr;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(256), 0, st, cnt, n_output_steps * 4);
    // Kernels with a work queue start each interval's systems longest first,
    // by their sub-step counts in the interval before (the first interval by
    // the caller's cost hints, if given); results do not depend on the order.
    const bool ordered = queue_kernel(net, opts);
    int32_t *d_steps = nullptr;
    const int32_t *d_index = nullptr;
    bool have_steps = false;
    if (ordered) {
        rc = reserve(c.out_steps, (size_t)Nsystems * 4);
        if (rc) return rc;
        d_steps = (int32_t *)c.out_steps.ptr;
        if (opts.cost_hint) {
            const int32_t *d_hint = opts.cost_hint;
            if (!opts.device_pointers) {
                rc = reserve(c.hint, (size_t)Nsystems * 4);
                if (rc) return rc;
                WIND_HIP(hipMemcpyAsync(c.hint.ptr, opts.cost_hint, (size_t)Nsystems * 4, hipMemcpyHostToDevice,
                                        st));
                d_hint = (const int32_t *)c.hint.ptr;
            }
            rc = order_by_hint(c.order, d_hint, Nsystems, st, &d_index);
            if (rc) return rc;
        }
    }
    SolverChoice sc;
    sc.solver = opts.solver;
    sc.ir_sweeps = opts.ir_sweeps;
    sc.pivot = (opts.flags & WIND_FLAG_PIVOT) ? 1 : 0;
    sc.dense_kernel = (opts.flags & WIND_FLAG_DENSE_LDS) ? 1 : 0;
    double tcur = tnow;
    WIND_HIP(hipEventRecord(c.out_events[0], st));
    for (int k = 1; k <= n_output_steps; k++) {
        float *slot = snap + eq_floats * (size_t)k;
        WIND_HIP(hipMemcpyAsync(slot, slot - eq_floats, eq_bytes, hipMemcpyDeviceToDevice, st));
        const float t0 = (float)tcur, t1 = (float)(tcur + dt);  // ctypes.c_float(tcur), c_float(tcur + dt)
        tcur += dt;
        if (t0 < t1) {
            IntegrateArgs a;
            a.tnow = t0;
            a.tend = t1;
            a.h0 = (t1 - t0) / (float)n_integration_steps;
            a.ABS = ABSOLUTE;
            a.REL = RELATIVE;
            a.controller = opts.controller;
            a.constants = d_const;
            a.equations = slot;
            a.nsys = Nsystems;
            a.const_stride = stride;
            a.max_attempts = opts.max_attempts > 0 ? opts.max_attempts : kDefaultMaxAttempts;
            if (have_steps) {
                rc = order_by_hint(c.order, d_steps, Nsystems, st, &d_index);
                if (rc) return rc;
            }
            a.steps_out = d_steps;
            a.status_out = nullptr;
            a.index = d_index;
            a.counters = cnt + 4 * (k - 1);
            rc = opts.method == WIND_SIE ? dispatch<WIND_SIE>(c.num_cu, net, a, opts.mode, opts.kernel, sc, st)
                                         : dispatch<WIND_RK2>(c.num_cu, net, a, opts.mode, opts.kernel, sc, st);
            if (rc) return rc;
            have_steps = ordered;
        }
        WIND_HIP(hipEventRecord(c.out_events[k], st));
    }
    std::vector<unsigned long long> host_cnt((size_t)n_output_steps * 4);
    WIND_HIP(hipMemcpyAsync(host_cnt.data(), cnt, host_cnt.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, st));
    const float *last = snap + eq_floats * (size_t)n_output_steps;
    if (opts.device_pointers) {
        WIND_HIP(hipMemcpyAsync(equations, last, eq_bytes, hipMemcpyDeviceToDevice, st));
        if (equations_over_time)
            WIND_HIP(hipMemcpyAsync(equations_over_time, snap, eq_bytes * (size_t)(n_output_steps + 1),
                                    hipMemcpyDeviceToDevice, st));
    } else {
        WIND_HIP(hipMemcpyAsync(equations, last, eq_bytes, hipMemcpyDeviceToHost, st));
        if (equations_over_time)
            WIND_HIP(hipMemcpyAsync(equations_over_time, snap, eq_bytes * (size_t)(n_output_steps + 1),
                                    hipMemcpyDeviceToHost, st));
    }
    WIND_HIP(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int k = 0; k < n_output_steps; k++) {
        if (nsteps) nsteps[k] = (int64_t)host_cnt[4 * (size_t)k + 1];
        total += (int64_t)host_cnt[4 * (size_t)k + 1];
        if (step_ms) {
            float ms = 0.0f;
            WIND_HIP(hipEventElapsedTime(&ms, c.out_events[k], c.out_events[k + 1]));
            step_ms[k] = ms;
        }
    }
    if (opts.total_steps) *opts.total_steps = total;
    return WIND_OK;
}

int wind_integrate(const wind_network *net, float tnow, float tend, int n_integration_steps,
                   const float *constants, float *equations, int Nsystems, int Neqn,
                   float ABSOLUTE, float RELATIVE, const wind_options *opts_in) {
    g_last_error.clear();
    wind_options opts;
    if (opts_in)
        opts = *opts_in;
    else
        wind_options_init(&opts);
    int rc = check_integrate_args(net, opts, Nsystems, Neqn, n_integration_steps);
    if (rc) return rc;
    // As in the reference, systems may share constants when the stride is
    // below the network's constant count (cuda/include/ode_gold.h:1 ships
    // NUM_CONST = 2 for a network that reads 10); the caller's array must
    // hold (Nsystems - 1) * stride + nconst floats.
    const int stride = opts.const_stride > 0 ? opts.const_stride : net->nconst;
    if (Nsystems == 0 || !(tnow < tend)) return WIND_OK;  // no attempt is made (common_gold.c:77)
    if (!constants || !equations) return fail(WIND_ERR_ARGUMENT, "NULL state or constants");

    // The call takes a slot of its own (stream, events, buffers): callers on
    // several threads only wait for each other where they share the matrix
    // scratch.
    Context &c = ctx();
    CallSlot *sl = nullptr;
    {
        std::lock_guard<std::mutex> lock(c.mu);
        rc = ensure_ready(c);
        if (!rc) rc = acquire_slot(c, &sl);
        if (rc) return rc;
    }
    SlotGuard release{c, sl};
    SolverChoice sc;
    sc.solver = opts.solver;
    sc.ir_sweeps = opts.ir_sweeps;
    sc.pivot = (opts.flags & WIND_FLAG_PIVOT) ? 1 : 0;
    sc.dense_kernel = (opts.flags & WIND_FLAG_DENSE_LDS) ? 1 : 0;
    size_t scratch_need = 0;
    rc = scratch_bytes_for(c.num_cu, net, opts.method, opts.mode, opts.kernel, sc, Nsystems, &scratch_need);
    if (rc) return rc;
    std::unique_lock<std::mutex> scratch_lock(g_scratch_mu, std::defer_lock);
    if (scratch_need) scratch_lock.lock();
    hipStream_t st = opts.stream ? (hipStream_t)opts.stream : sl->stream;

    const size_t eq_bytes = (size_t)Nsystems * Neqn * sizeof(float);
    // Reads of constants reach (Nsystems - 1) * stride + nconst floats.
    const size_t c_floats = (size_t)(Nsystems - 1) * stride + net->nconst;
    const size_t c_bytes = c_floats * sizeof(float);

    const float *d_const = constants;
    float *d_eq = equations;
    int32_t *d_steps = opts.steps_out;
    int32_t *d_status = opts.status_out;
    if (!opts.device_pointers) {
        rc = reserve(sl->constants, c_bytes);
        if (!rc) rc = reserve(sl->equations, eq_bytes);
        if (!rc && opts.steps_out) rc = reserve(sl->steps, (size_t)Nsystems * 4);
        if (!rc && opts.status_out) rc = reserve(sl->status, (size_t)Nsystems * 4);
        if (rc) return rc;
        WIND_HIP(hipMemcpyAsync(sl->constants.ptr, constants, c_bytes, hipMemcpyHostToDevice, st));
        WIND_HIP(hipMemcpyAsync(sl->equations.ptr, equations, eq_bytes, hipMemcpyHostToDevice, st));
        d_const = (const float *)sl->constants.ptr;
        d_eq = (float *)sl->equations.ptr;
        d_steps = opts.steps_out ? (int32_t *)sl->steps.ptr : nullptr;
        d_status = opts.status_out ? (int32_t *)sl->status.ptr : nullptr;
    }
    rc = reserve(sl->counters, 4 * sizeof(unsigned long long));
    if (rc) return rc;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long *)sl->counters.ptr, 4);

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
    a.counters = (unsigned long long *)sl->counters.ptr;

    WIND_HIP(hipEventRecord(sl->ev0, st));
    // Start order for the kernels that take systems from a queue: from the
    // caller's cost hints, or (fast mode, compiled-in networks, batches
    // larger than the device runs at once) from an estimate of each
    // system's work made from its initial state.
    // Lane kernels of the compiled-in networks also get the systems packed
    // in that order (records), so that their lanes read contiguously.
    PackArgs pk;
    const bool pack = net->builtin == WIND_NET_NR_TEST || net->builtin == WIND_NET_KATZ96 ||
                      net->builtin == WIND_NET_HHE6;
    if (pack && (opts.cost_hint || auto_order(net, opts, Nsystems)) && queue_kernel(net, opts)) {
        pk.eq = a.equations;
        pk.cst = a.constants;
        pk.N = Neqn;
        pk.NC = net->nconst;
        pk.stride = a.const_stride;
        pk.R = lane_record_words(Neqn, net->nconst);
        rc = reserve(sl->records, (size_t)Nsystems * pk.R * sizeof(float));
        if (rc) return rc;
        pk.rec = (float *)sl->records.ptr;
    }
    if (opts.cost_hint && queue_kernel(net, opts)) {
        const int32_t *d_hint = opts.cost_hint;
        if (!opts.device_pointers) {
            rc = reserve(sl->hint, (size_t)Nsystems * 4);
            if (rc) return rc;
            WIND_HIP(hipMemcpyAsync(sl->hint.ptr, opts.cost_hint, (size_t)Nsystems * 4, hipMemcpyHostToDevice, st));
            d_hint = (const int32_t *)sl->hint.ptr;
        }
        rc = order_by_hint(sl->order, d_hint, Nsystems, st, &a.index, &pk);
        if (rc) return rc;
    } else if (auto_order(net, opts, Nsystems)) {
        rc = reserve(sl->hint, (size_t)Nsystems * 4);
        if (!rc) rc = reserve(sl->activity, ((size_t)Nsystems + 64) * 4);
        if (rc) return rc;
        hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long *)sl->activity.ptr, 32);
        rc = activity_hints(c.num_cu, net, a, opts.method, (int32_t *)sl->hint.ptr, (unsigned *)sl->activity.ptr, st);
        if (!rc) rc = order_by_hint(sl->order, (const int32_t *)sl->hint.ptr, Nsystems, st, &a.index, &pk);
        if (rc) return rc;
    }
    a.records = pk.rec;
    std::vector<TraceCopy> trace_back;
    if (opts.trace) {
        rc = setup_trace(sl->trace, opts, Nsystems, Neqn, &a.trace, &trace_back, st);
        if (rc) return rc;
    }
    rc = opts.method == WIND_SIE ? dispatch<WIND_SIE>(c.num_cu, net, a, opts.mode, opts.kernel, sc, st)
                                 : dispatch<WIND_RK2>(c.num_cu, net, a, opts.mode, opts.kernel, sc, st);
    if (rc) return rc;
    WIND_HIP(hipEventRecord(sl->ev1, st));

    unsigned long long total = 0;
    WIND_HIP(hipMemcpyAsync(&total, a.counters + 1, sizeof(total), hipMemcpyDeviceToHost, st));
    for (const TraceCopy &tc : trace_back)
        WIND_HIP(hipMemcpyAsync(tc.dst, tc.src, tc.bytes, hipMemcpyDeviceToHost, st));
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
    if (opts.kernel_ms) WIND_HIP(hipEventElapsedTime(opts.kernel_ms, sl->ev0, sl->ev1));
    if (opts.total_steps) *opts.total_steps = (int64_t)total;
    return WIND_OK;
}

}  // extern "C"

static bool auto_order(const wind_network *net, const wind_options &opts, int Nsystems) {
    return opts.mode == WIND_MODE_FAST && opts.kernel != WIND_KERNEL_BLOCK && Nsystems >= kAutoOrderMin &&
           (net->builtin == WIND_NET_NR_TEST || net->builtin == WIND_NET_KATZ96 || net->builtin == WIND_NET_HHE6);
}

// ---------------------------------------------------------------------------
// wind_integrate_noalloc: wind_integrate for callers that cannot allocate.
// Everything the kernels need beyond the caller's inputs and outputs (the
// work-queue head and sub-step total, and the start-order buffers) lives in
// one statically sized device array, `scratch`.  Its device address is taken
// once per device with hipGetSymbolAddress (the array's name is only a
// host-side stand-in) and used for every kernel argument.  The array is
// shared by every call on a device and keeps the last call's contents; the
// calls of one process are serialised by g_noalloc_mu, so they never overlap
// on it.
// ---------------------------------------------------------------------------
namespace {

// 32-bit words of the scratch layout.
constexpr int kNoAllocCounterWords = 64;               // 4 x u64 counters, padded to 256 B
constexpr int kNoAllocCountWords = kHintCountWords;     // start-order class counts
constexpr size_t kNoAllocWords = (size_t)kNoAllocCounterWords + kNoAllocCountWords +
                                 (64 + (size_t)WIND_NOALLOC_MAX_SYSTEMS) +  // activity estimates
                                 (size_t)WIND_NOALLOC_MAX_SYSTEMS +         // hints
                                 (size_t)WIND_NOALLOC_MAX_SYSTEMS;          // start order
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

struct NoAllocDevice {
    bool ready = false;
    uint32_t *scratch = nullptr;  // device address of `scratch` on this device
    int num_cu = 0;
};
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kMaxDevices];

// Once per device, with that device current: the array's device address
// and the CU count.
int noalloc_setup(int dev, NoAllocDevice &d) {
    if (d.ready) return WIND_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(WIND_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    WIND_HIP(hipDeviceGetAttribute(&d.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    d.scratch = (uint32_t *)ptr;
    d.ready = true;
    return WIND_OK;
}

}  // namespace

extern "C" int wind_integrate_noalloc(const wind_network *net, float tnow, float tend, int n_integration_steps,
                                      const float *constants, float *equations, int Nsystems, int Neqn,
                                      float ABSOLUTE, float RELATIVE, const wind_options *opts_in) {
    g_last_error.clear();
    wind_options opts;
    if (opts_in)
        opts = *opts_in;
    else
        wind_options_init(&opts);
    int rc = check_integrate_args(net, opts, Nsystems, Neqn, n_integration_steps);
    if (rc) return rc;
    if (!opts.device_pointers)
        return fail(WIND_ERR_ARGUMENT, "wind_integrate_noalloc takes device pointers only (device_pointers = 1)");
    if (opts.kernel_ms) return fail(WIND_ERR_ARGUMENT, "wind_integrate_noalloc does not time its kernels (kernel_ms)");
    if (Nsystems > WIND_NOALLOC_MAX_SYSTEMS)
        return fail(WIND_ERR_ARGUMENT, "wind_integrate_noalloc: Nsystems " + std::to_string(Nsystems) +
                                           " exceeds its scratch maximum of " +
                                           std::to_string(WIND_NOALLOC_MAX_SYSTEMS) + " systems");

    const int stride = opts.const_stride > 0 ? opts.const_stride : net->nconst;
    if (Nsystems == 0 || !(tnow < tend)) return WIND_OK;
    if (!constants || !equations) return fail(WIND_ERR_ARGUMENT, "NULL state or constants");

    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    int dev = 0;
    WIND_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(WIND_ERR_DEVICE, "device index out of range");
    NoAllocDevice &d = g_noalloc[dev];
    rc = noalloc_setup(dev, d);
    if (rc) return rc;
    {
        SolverChoice sc0;
        sc0.pivot = (opts.flags & WIND_FLAG_PIVOT) ? 1 : 0;
        size_t need = 0;
        rc = scratch_bytes_for(d.num_cu, net, opts.method, opts.mode, opts.kernel, sc0, Nsystems, &need);
        if (rc) return rc;
        if (need)
            return fail(WIND_ERR_UNSUPPORTED,
                        "wind_integrate_noalloc: this network's matrices do not fit in LDS (they need an allocation)");
    }
    hipStream_t st = (hipStream_t)opts.stream;  // NULL: the device's null stream
    uint32_t *base = d.scratch;
    unsigned long long *counters = (unsigned long long *)base;
    unsigned *counts = base + kNoAllocCounterWords;
    unsigned *activity = counts + kNoAllocCountWords;
    int32_t *hint = (int32_t *)(activity + 64 + WIND_NOALLOC_MAX_SYSTEMS);
    int32_t *order = hint + WIND_NOALLOC_MAX_SYSTEMS;
    // The counters start at zero (a fill kernel, not a memset).
    wind_zero_counters(counters, 4, st);

    IntegrateArgs a;
    a.tnow = tnow;
    a.tend = tend;
    a.h0 = (tend - tnow) / (float)n_integration_steps;
    a.ABS = ABSOLUTE;
    a.REL = RELATIVE;
    a.controller = opts.controller;
    a.constants = constants;
    a.equations = equations;
    a.nsys = Nsystems;
    a.const_stride = stride;
    a.max_attempts = opts.max_attempts > 0 ? opts.max_attempts : kDefaultMaxAttempts;
    a.steps_out = opts.steps_out;
    a.status_out = opts.status_out;
    a.index = nullptr;
    a.counters = counters;
    // The same start orders as wind_integrate, built in the scratch array.
    if (opts.cost_hint && queue_kernel(net, opts)) {
        wind_order_by_hint(opts.cost_hint, Nsystems, order, counts, st);
        WIND_HIP(hipGetLastError());
        a.index = order;
    } else if (auto_order(net, opts, Nsystems)) {
        wind_zero_counters((unsigned long long *)activity, 32, st);
        rc = activity_hints(d.num_cu, net, a, opts.method, hint, activity, st);
        if (rc) return rc;
        wind_order_by_hint(hint, Nsystems, order, counts, st);
        WIND_HIP(hipGetLastError());
        a.index = order;
    }
    if (opts.trace) {
        // device pointers only: the trace arrays are the caller's, nothing is staged
        Pool unused;
        std::vector<TraceCopy> back;
        rc = setup_trace(unused, opts, Nsystems, Neqn, &a.trace, &back, st);
        if (rc) return rc;
    }
    SolverChoice sc;
    sc.solver = opts.solver;
    sc.ir_sweeps = opts.ir_sweeps;
    sc.pivot = (opts.flags & WIND_FLAG_PIVOT) ? 1 : 0;
    sc.dense_kernel = (opts.flags & WIND_FLAG_DENSE_LDS) ? 1 : 0;
    rc = opts.method == WIND_SIE ? dispatch<WIND_SIE>(d.num_cu, net, a, opts.mode, opts.kernel, sc, st)
                                 : dispatch<WIND_RK2>(d.num_cu, net, a, opts.mode, opts.kernel, sc, st);
    if (rc) return rc;
    unsigned long long total = 0;
    WIND_HIP(hipMemcpyAsync(&total, counters + 1, sizeof(total), hipMemcpyDeviceToHost, st));
    WIND_HIP(hipStreamSynchronize(st));
    if (opts.total_steps) *opts.total_steps = (int64_t)total;
    return WIND_OK;
}

namespace {


int launch_bucket(const Context &c, const wind_network *net, const IntegrateArgs &a, const wind_options &o,
                  const SolverChoice &sc, hipStream_t st) {
    return o.method == WIND_SIE ? dispatch<WIND_SIE>(c.num_cu, net, a, o.mode, WIND_KERNEL_AUTO, sc, st)
                                : dispatch<WIND_RK2>(c.num_cu, net, a, o.mode, WIND_KERNEL_AUTO, sc, st);
}

IntegrateArgs bucket_args(const wind_bucket &b, float tnow, float tend, int nfix, float ABS, float REL,
                          const wind_options &o, unsigned long long *counters) {
    IntegrateArgs a;
    a.tnow = tnow;
    a.tend = tend;
    a.h0 = (tend - tnow) / (float)nfix;
    a.ABS = ABS;
    a.REL = REL;
    a.controller = o.controller;
    a.constants = b.constants;
    a.equations = b.equations;
    a.nsys = b.nsys;
    a.const_stride = b.const_stride > 0 ? b.const_stride : b.net->nconst;
    a.max_attempts = o.max_attempts > 0 ? o.max_attempts : kDefaultMaxAttempts;
    a.steps_out = b.steps_out;
    a.status_out = b.status_out;
    a.index = nullptr;
    a.counters = counters;
    return a;
}

// Candidate kernel shapes for a bucket; the first is the default.
std::vector<SolverChoice> candidates(const wind_network *net, const wind_options &o) {
    std::vector<SolverChoice> v;
    SolverChoice base;
    base.solver = o.solver;
    base.ir_sweeps = o.ir_sweeps;
    base.pivot = (o.flags & WIND_FLAG_PIVOT) ? 1 : 0;
    base.dense_kernel = (o.flags & WIND_FLAG_DENSE_LDS) ? 1 : 0;
    v.push_back(base);
    if (net->builtin != WIND_NET_MASS_ACTION || o.mode != WIND_MODE_FAST || base.pivot) return v;
    const int NP = (net->N + 15) & ~15;
    if (NP > 192) return v;
    v.clear();
    // the tile products of the dense solver: FP32 MFMA unless the caller names one
    const int sv = o.solver == WIND_SOLVER_AUTO ? WIND_SOLVER_MFMA_F32 : o.solver;
    const int blocks[4] = {64, 128, 256, 512};
    // Where the register-resident solver applies (SIE) it is the only candidate:
    // the LDS solver rounds differently, and a bucket's results must not
    // depend on the tuning (every shape of one solver gives the same bits).
    if (o.method == WIND_SIE) {
        SolverChoice sc = base;
        sc.solver = sv;
        sc.dense_kernel = 2;
        if (base.dense_kernel != 1 && dense_reg_eligible(net->ma, sc)) {
            v.push_back(sc);
            return v;
        }
    }
    {
        for (int blk : blocks) {
            if (NP > 64 && blk == 64) continue;
            if (NP <= 32 && blk > 128) continue;
            SolverChoice sc;
            sc.solver = sv;
            sc.ir_sweeps = o.ir_sweeps;
            sc.pivot = base.pivot;
            sc.block = blk;
            sc.dense_kernel = 1;
            v.push_back(sc);
        }
    }
    if (v.empty()) v.push_back(base);
    return v;
}

int tune_bucket(Context &c, const wind_bucket &b, float tnow, float tend, int nfix, float ABS, float REL,
                const wind_options &o, hipStream_t st, TuneVal *out) {
    TuneKey key{b.net, o.method, o.mode, o.controller, o.solver, o.ir_sweeps,
                ((o.flags & WIND_FLAG_PIVOT) ? 1 : 0) | ((o.flags & WIND_FLAG_DENSE_LDS) ? 2 : 0)};
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
        *out = it->second;
        return WIND_OK;
    }
    const int N = b.net->N;
    // samples that fill the device several times over (the register-resident
    // dense solver runs three or four systems per CU)
    const int ns = std::min(b.nsys, N <= 16 ? 16384 : (N <= 64 ? 4096 : 12 * c.num_cu));
    const size_t bytes = (size_t)ns * N * sizeof(float);
    int rc = reserve(c.tune_state, bytes);
    if (!rc) rc = reserve(c.counters, 4 * sizeof(unsigned long long));
    if (rc) return rc;
    wind_bucket sb = b;
    sb.nsys = ns;
    sb.equations = (float *)c.tune_state.ptr;
    sb.steps_out = nullptr;
    sb.status_out = nullptr;
    std::vector<SolverChoice> cand = candidates(b.net, o);
    TuneVal best;
    best.sec_per_sys = 1e300;
    for (const SolverChoice &sc : cand) {
        WIND_HIP(hipMemcpyAsync(c.tune_state.ptr, b.equations, bytes, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long *)c.counters.ptr, 4);
        IntegrateArgs a = bucket_args(sb, tnow, tend, nfix, ABS, REL, o, (unsigned long long *)c.counters.ptr);
        WIND_HIP(hipEventRecord(c.ev0, st));
        rc = launch_bucket(c, b.net, a, o, sc, st);
        if (rc) return rc;
        WIND_HIP(hipEventRecord(c.ev1, st));
        WIND_HIP(hipEventSynchronize(c.ev1));
        float ms = 0.0f;
        WIND_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        const double sps = ms * 1e-3 / ns;
        if (sps < best.sec_per_sys) {
            best.sec_per_sys = sps;
            best.sc = sc;
        }
    }
    g_tuned[key] = best;
    *out = best;
    return WIND_OK;
}

}  // namespace

extern "C" int wind_integrate_batch(const wind_bucket *buckets, int nb, float tnow, float tend, int nfix,
                                    float ABSOLUTE, float RELATIVE, const wind_options *opts_in) {
    g_last_error.clear();
    wind_options o;
    if (opts_in)
        o = *opts_in;
    else
        wind_options_init(&o);
    if (o.total_steps) *o.total_steps = 0;
    if (o.kernel_ms) *o.kernel_ms = 0.0f;
    if (nb < 0 || (nb > 0 && !buckets)) return fail(WIND_ERR_ARGUMENT, "bad bucket list");
    if (o.method != WIND_SIE && o.method != WIND_RK2) return fail(WIND_ERR_ARGUMENT, "bad method");
    if (o.mode != WIND_MODE_EXACT && o.mode != WIND_MODE_FAST) return fail(WIND_ERR_ARGUMENT, "bad mode");
    if (o.controller != WIND_CTRL_REFERENCE && o.controller != WIND_CTRL_ADAPTIVE)
        return fail(WIND_ERR_ARGUMENT, "bad controller");
    if (o.solver < WIND_SOLVER_AUTO || o.solver > WIND_SOLVER_MFMA_BF16X3) return fail(WIND_ERR_ARGUMENT, "bad solver");
    if (o.ir_sweeps < 0 || o.ir_sweeps > 4) return fail(WIND_ERR_ARGUMENT, "ir_sweeps must be 0-4");
    if (nfix < 0) return fail(WIND_ERR_ARGUMENT, "n_integration_steps < 0");
    if (o.trace) return fail(WIND_ERR_ARGUMENT, "trace: supported by wind_integrate only");
    if ((o.flags & WIND_FLAG_PIVOT) && o.mode != WIND_MODE_FAST)
        return fail(WIND_ERR_ARGUMENT, "WIND_FLAG_PIVOT applies to fast mode");
    for (int i = 0; i < nb; i++) {
        const wind_bucket &b = buckets[i];
        if (!b.net || b.nsys < 0 || (b.nsys > 0 && (!b.constants || !b.equations)))
            return fail(WIND_ERR_ARGUMENT, "bucket " + std::to_string(i) + ": bad network, size or pointers");
    }
    if (nb == 0 || !(tnow < tend)) return WIND_OK;
    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    hipStream_t st = o.stream ? (hipStream_t)o.stream : c.stream;

    // Tune each bucket (cached).
    std::vector<TuneVal> tv(nb);
    for (int i = 0; i < nb; i++) {
        if (buckets[i].nsys == 0) continue;
        if (o.flags & WIND_FLAG_NO_AUTOTUNE) {
            tv[i].sc = candidates(buckets[i].net, o)[0];
            tv[i].sec_per_sys = 1.0;
        } else {
            rc = tune_bucket(c, buckets[i], tnow, tend, nfix, ABSOLUTE, RELATIVE, o, st, &tv[i]);
            if (rc) return rc;
        }
    }
    // Every bucket gets a full-device persistent grid, and buckets are
    // launched on their own streams in decreasing cost per system.  The most
    // expensive bucket starts on every CU; as its queue drains, its workgroups
    // exit and the dispatcher backfills the freed CUs with the next buckets'
    // workgroups, so no CU idles while any bucket still has systems.
    std::vector<SolverChoice> sc(nb);
    for (int i = 0; i < nb; i++) sc[i] = tv[i].sc;
    std::vector<int> order;
    for (int i = 0; i < nb; i++)
        if (buckets[i].nsys > 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return tv[x].sec_per_sys > tv[y].sec_per_sys; });
    // Buckets whose matrices exceed LDS share the process-wide scratch: it is
    // grown here, before anything is enqueued or captured, and those buckets
    // run one after another on the stream of the first of them.
    std::vector<int> stream_of(nb);
    {
        size_t need = 0;
        int first = -1;
        for (int i : order) {
            size_t b = 0;
            rc = scratch_bytes_for(c.num_cu, buckets[i].net, o.method, o.mode, WIND_KERNEL_AUTO, sc[i],
                                   buckets[i].nsys, &b);
            if (rc) return rc;
            if (b > 0 && first < 0) first = i;
            stream_of[i] = b > 0 ? first : i;
            need = std::max(need, b);
        }
        if (need > 0) {
            rc = mass_action_scratch_reserve(need);
            if (rc) return rc;
        }
    }

    while ((int)c.bucket_streams.size() < nb) {
        hipStream_t s2;
        hipEvent_t e2;
        WIND_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        WIND_HIP(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
        c.bucket_streams.push_back(s2);
        c.bucket_events.push_back(e2);
    }
    if (!c.fork) WIND_HIP(hipEventCreateWithFlags(&c.fork, hipEventDisableTiming));
    rc = reserve(c.batch_counters, (size_t)nb * 4 * sizeof(unsigned long long));
    if (rc) return rc;
    unsigned long long *cnt = (unsigned long long *)c.batch_counters.ptr;
    // Start orders from the buckets' cost hints: buffers reserved here, before
    // anything is enqueued or captured; the ordering kernels run on each
    // bucket's stream ahead of its integration kernel.
    if (c.bucket_order.size() < (size_t)nb) c.bucket_order.resize(nb);
    std::vector<int32_t *> order_index(nb, nullptr);
    std::vector<unsigned *> order_counts(nb, nullptr);
    for (int i : order) {
        if (!buckets[i].cost_hint || !queue_kernel(buckets[i].net, o)) continue;
        const size_t index_bytes = ((size_t)buckets[i].nsys * 4 + 255) & ~(size_t)255;
        rc = reserve(c.bucket_order[i], index_bytes + kHintClasses * sizeof(unsigned));
        if (rc) return rc;
        order_index[i] = (int32_t *)c.bucket_order[i].ptr;
        order_counts[i] = (unsigned *)((char *)c.bucket_order[i].ptr + index_bytes);
    }

    auto enqueue = [&]() -> int {
        hipLaunchKernelGGL(zero_u64, dim3(1), dim3(256), 0, st, cnt, nb * 4);
        WIND_HIP(hipEventRecord(c.fork, st));
        for (int i : order) {
            hipStream_t bs = c.bucket_streams[stream_of[i]];
            WIND_HIP(hipStreamWaitEvent(bs, c.fork, 0));
            IntegrateArgs a = bucket_args(buckets[i], tnow, tend, nfix, ABSOLUTE, RELATIVE, o, cnt + 4 * i);
            if (order_index[i]) {
                wind_order_by_hint(buckets[i].cost_hint, buckets[i].nsys, order_index[i], order_counts[i], bs);
                a.index = order_index[i];
            }
            int r = launch_bucket(c, buckets[i].net, a, o, sc[i], bs);
            if (r) return r;
            WIND_HIP(hipEventRecord(c.bucket_events[i], bs));
            WIND_HIP(hipStreamWaitEvent(st, c.bucket_events[i], 0));
        }
        return WIND_OK;
    };

    WIND_HIP(hipEventRecord(c.ev0, st));
    if (o.flags & WIND_FLAG_GRAPH) {
        std::vector<uint64_t> key;
        auto push = [&](uint64_t v) { key.push_back(v); };
        push((uint64_t)(uintptr_t)st);
        push(((uint64_t)__float_as_uint_host(tnow) << 32) | __float_as_uint_host(tend));
        push(((uint64_t)__float_as_uint_host(ABSOLUTE) << 32) | __float_as_uint_host(RELATIVE));
        push(((uint64_t)(uint32_t)nfix << 32) | (uint32_t)o.method | ((uint64_t)o.mode << 8) |
             ((uint64_t)o.controller << 16));
        push((uint64_t)o.max_attempts);
        for (int i = 0; i < nb; i++) {
            const wind_bucket &b = buckets[i];
            push((uint64_t)(uintptr_t)b.net);
            push((uint64_t)(uintptr_t)b.constants);
            push((uint64_t)(uintptr_t)b.equations);
            push((uint64_t)(uintptr_t)b.steps_out);
            push((uint64_t)(uintptr_t)b.status_out);
            push((uint64_t)(uintptr_t)b.cost_hint);
            push((uint64_t)(uintptr_t)order_index[i]);
            push(((uint64_t)(uint32_t)b.nsys << 32) | (uint32_t)b.const_stride);
            push(((uint64_t)(uint32_t)sc[i].solver << 48) | ((uint64_t)(uint32_t)sc[i].block << 32) |
                 ((uint64_t)(uint32_t)sc[i].ir_sweeps << 24) | ((uint64_t)(uint32_t)sc[i].pivot << 20) |
                 ((uint64_t)(uint32_t)sc[i].dense_kernel << 18) | (uint32_t)sc[i].grid);
            push((uint64_t)__float_as_uint_host(sc[i].share));
        }
        push((uint64_t)(uintptr_t)cnt);
        // Buckets beyond LDS take the matrix scratch's address as a kernel
        // argument: a graph captured before the scratch grew (and was freed)
        // must not be replayed.
        push((uint64_t)(uintptr_t)mass_action_scratch_ptr());
        if (!c.graph_exec || key != c.graph_key) {
            if (c.graph_exec) {
                (void)hipGraphExecDestroy(c.graph_exec);
                c.graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            // Thread-local capture: other threads of the process (a server's
            // scheduler) may keep calling the runtime meanwhile.
            WIND_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = enqueue();
            hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : fail(WIND_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
            }
            WIND_HIP(hipGraphInstantiate(&c.graph_exec, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            c.graph_key = key;
        }
        WIND_HIP(hipGraphLaunch(c.graph_exec, st));
    } else {
        rc = enqueue();
        if (rc) return rc;
    }
    WIND_HIP(hipEventRecord(c.ev1, st));
    std::vector<unsigned long long> host(nb * 4);
    WIND_HIP(hipMemcpyAsync(host.data(), cnt, host.size() * 8, hipMemcpyDeviceToHost, st));
    WIND_HIP(hipStreamSynchronize(st));
    if (o.kernel_ms) WIND_HIP(hipEventElapsedTime(o.kernel_ms, c.ev0, c.ev1));
    long long total = 0;
    for (int i = 0; i < nb; i++) total += (long long)host[4 * i + 1];
    if (o.total_steps) *o.total_steps = total;
    return WIND_OK;
}

void wind_order_by_hint(const int32_t *hint, int n, int32_t *index, unsigned *count, hipStream_t st,
                        const PackArgs *pk) {
    static_assert(kHintCountWords == kHintClasses, "one count per class");
    const int blocks = std::min(kHintBlocks, (n + 255) / 256);
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long *)count, kHintClasses / 2);
    hipLaunchKernelGGL(hint_count, dim3(blocks), dim3(256), 0, st, hint, n, count);
    hipLaunchKernelGGL(hint_starts, dim3(1), dim3(64), 0, st, count);
    hipLaunchKernelGGL(hint_place, dim3(blocks), dim3(256), 0, st, hint, n, count, index, pk ? *pk : PackArgs{});
}

// The lane kernels of compiled-in and generated networks and the dense
// solver of mass-action networks (fast mode) take systems from a queue.
bool wind_queue_kernel(const wind_network *net, int mode, int kernel) {
    return kernel != WIND_KERNEL_BLOCK &&
           (net->builtin == WIND_NET_NR_TEST || net->builtin == WIND_NET_KATZ96 || net->builtin == WIND_NET_HHE6 ||
            net->builtin == WIND_NET_PLUGIN || (net->builtin == WIND_NET_MASS_ACTION && mode == WIND_MODE_FAST));
}

int wind_launch_async(const wind_network *net, const IntegrateArgs &a, int method, int mode, int kernel,
                      const SolverChoice &sc, hipStream_t st) {
    Context &c = ctx();
    {
        std::lock_guard<std::mutex> lock(c.mu);
        int rc = ensure_ready(c);
        if (rc) return rc;
    }
    size_t need = 0;
    const int rc = scratch_bytes_for(c.num_cu, net, method, mode, kernel, sc, a.nsys, &need);
    if (rc) return rc;
    if (need)
        return fail(WIND_ERR_UNSUPPORTED, "this network's matrices do not fit in LDS (it needs the shared scratch buffer)");
    return method == WIND_SIE ? dispatch<WIND_SIE>(c.num_cu, net, a, mode, kernel, sc, st)
                              : dispatch<WIND_RK2>(c.num_cu, net, a, mode, kernel, sc, st);
}

int wind_device_cus(int *num_cu) {
    Context &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    int rc = ensure_ready(c);
    if (rc) return rc;
    *num_cu = c.num_cu;
    return WIND_OK;
}

void wind_zero_counters(unsigned long long *p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, p, n);
}
