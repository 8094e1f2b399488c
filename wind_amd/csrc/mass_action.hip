// Host side of the mass-action networks: gather-list construction, device
// upload, and kernel launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "kernels_dense.h"
#include "kernels_dense_reg.h"
#include "kernels_mass_action.h"
#include "wind_internal.h"

using namespace wind;

namespace {

int upload(void **dst, const void *src, size_t bytes) {
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    if (src) {
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    }
    return WIND_OK;
}

struct Scratch {
    float *ptr = nullptr;
    size_t bytes = 0;
};
Scratch g_scratch;  // per-process matrix scratch; callers hold the context lock

}  // namespace

int mass_action_build(wind_network *n, int N, int R, const int32_t *rx, const float *k) {
    n->rx.assign(rx, rx + 4 * (size_t)R);
    n->k.assign(k, k + R);
    // Contribution lists in the order of a serial sweep over reactions.
    std::vector<std::vector<int32_t>> flist(N);
    std::map<int, std::vector<int32_t>> jlist;
    for (int r = 0; r < R; r++) {
        const int32_t *x = rx + 4 * r;
        for (int slot = 0; slot < 4; slot++) {
            const int neg = slot < 2 ? 1 : 0;
            flist[x[slot]].push_back(r * 2 + neg);
            jlist[x[0] * N + x[slot]].push_back(((r * 2 + 0) * 2) + neg);
            jlist[x[1] * N + x[slot]].push_back(((r * 2 + 1) * 2) + neg);
        }
    }
    std::vector<int32_t> f_ptr(N + 1, 0), f_ent;
    for (int i = 0; i < N; i++) {
        f_ent.insert(f_ent.end(), flist[i].begin(), flist[i].end());
        f_ptr[i + 1] = (int32_t)f_ent.size();
    }
    std::vector<int32_t> j_pos, j_ptr(1, 0), j_ent;
    for (auto &kv : jlist) {
        j_pos.push_back(kv.first);
        j_ent.insert(j_ent.end(), kv.second.begin(), kv.second.end());
        j_ptr.push_back((int32_t)j_ent.size());
    }
    std::vector<std::vector<int32_t>> rows(N);
    for (size_t z = 0; z < j_pos.size(); z++) rows[j_pos[z] % N].push_back((int32_t)z);
    std::vector<int32_t> r_ptr(N + 1, 0), r_z;
    for (int i = 0; i < N; i++) {
        r_z.insert(r_z.end(), rows[i].begin(), rows[i].end());
        r_ptr[i + 1] = (int32_t)r_z.size();
    }
    // Assembly of the register-resident dense solver.  The matrix elements
    // (row, col) with Jacobian terms live, once summed, in a compact LDS
    // buffer (the others are 1 on the diagonal, 0 elsewhere), ordered by
    // 16x16 tile (row-major over tiles), then by lane
    // l = 16 ((row % 16) / 4) + col % 16 of the MFMA accumulator layout, then
    // by that lane's accumulator register q = row % 4.  Per tile and lane,
    // sc_tile[tile * 64 + l] = (offset of the lane's first element) << 4 |
    // (mask of the q present), so that the owner lane finds element q at
    // that offset plus the present q below it.  An element is its terms in
    // the order dense_eval applies them (the row's reactions in sweep order,
    // the reaction's first reactant before its second), summed from 1 on the
    // diagonal; the elements are dealt to the workgroup's 256
    // threads, most terms first to the least-loaded thread, and a thread's
    // stream is its elements' terms, one word each:
    //   (LDS float index of the term's h dJ value) << 16 | (buffer offset of
    //   the element) << 3 | (element on the diagonal) << 2 | (element's last
    //   term) << 1 | (reactant row: the term adds)
    // Streams are padded with zero words (which store nothing) to the
    // longest, rounded up to kDenseRegBatch words, and interleaved: word j
    // of thread t at sc_ent[j * 256 + t].
    const int NT = (N + 15) / 16;
    std::vector<int32_t> sc_ent;
    std::vector<uint32_t> sc_tile;
    int sc_len = 0, sc_cap = 0;
    if (dense_reg_tiles(NT)) {
        const DenseRegOffsets lo = dense_reg_offsets(NT, R, 3, 0);
        bool ok = lo.hdb + R < 65536;
        std::map<std::pair<int, int>, std::vector<uint32_t>> elem;  // (row, col) -> words without last
        for (int i = 0; i < N && ok; i++)
            for (int32_t e : flist[i]) {
                const int r = e >> 1, neg = e & 1;
                const int32_t *x = rx + 4 * r;
                for (int which = 0; which < 2; which++) {
                    const uint32_t hd = (uint32_t)((which ? lo.hdb : lo.hda) + r);
                    elem[{i, x[which]}].push_back((hd << 16) | ((x[which] == i ? 1u : 0u) << 2) | (uint32_t)neg);
                }
            }
        // the q present per (tile, lane), then the offsets
        std::vector<uint32_t> mask4((size_t)NT * NT * 64, 0);
        auto slot = [&](int row, int col, int *tl, int *q) {
            *tl = ((row >> 4) * NT + (col >> 4)) * 64 + 16 * ((row & 15) >> 2) + (col & 15);
            *q = row & 3;
        };
        for (auto &kv : elem) {
            int tl, q;
            slot(kv.first.first, kv.first.second, &tl, &q);
            mask4[tl] |= 1u << q;
        }
        sc_tile.resize(mask4.size());
        uint32_t nnz = 0;
        for (size_t u = 0; u < mask4.size(); u++) {
            sc_tile[u] = nnz << 4 | mask4[u];
            nnz += (uint32_t)__builtin_popcount(mask4[u]);
        }
        // (+1: an owner lane reads one place past its elements, unused)
        sc_cap = (int)((nnz + 1 + 3) & ~3u);
        ok = ok && sc_cap < 8192;
        std::vector<std::pair<int, std::vector<uint32_t>>> items;  // (count, words)
        for (auto &kv : elem) {
            int tl, q;
            slot(kv.first.first, kv.first.second, &tl, &q);
            const uint32_t dest = (sc_tile[tl] >> 4) + (uint32_t)__builtin_popcount(mask4[tl] & ((1u << q) - 1));
            std::vector<uint32_t> w = kv.second;
            for (auto &x : w) x |= dest << 3;
            w.back() |= 2u;
            items.push_back({(int)w.size(), std::move(w)});
        }
        std::stable_sort(items.begin(), items.end(),
                         [](const auto &p1, const auto &p2) { return p1.first > p2.first; });
        std::vector<std::vector<uint32_t>> per(256);
        for (auto &it : items) {
            size_t th = 0;
            for (size_t u = 1; u < per.size(); u++)
                if (per[u].size() < per[th].size()) th = u;
            per[th].insert(per[th].end(), it.second.begin(), it.second.end());
        }
        size_t len = kDenseRegBatch;
        for (auto &p : per) len = std::max(len, (p.size() + kDenseRegBatch - 1) / kDenseRegBatch * kDenseRegBatch);
        sc_ent.assign(len * 256, 0);
        for (int th = 0; th < 256; th++)
            for (size_t j = 0; j < per[th].size(); j++) sc_ent[j * 256 + th] = (int32_t)per[th][j];
        sc_len = (int)len;
        if (!ok) {
            sc_ent.clear();
            sc_tile.clear();
            sc_len = sc_cap = 0;
        }
    }
    MassActionDevice &d = n->ma;
    d.N = N;
    d.R = R;
    d.nnz = (int)j_pos.size();
    int rc = upload((void **)&d.rx, rx, sizeof(int32_t) * 4 * R);
    if (!rc) rc = upload((void **)&d.k, k, sizeof(float) * R);
    if (!rc) rc = upload((void **)&d.f_ptr, f_ptr.data(), sizeof(int32_t) * f_ptr.size());
    if (!rc) rc = upload((void **)&d.f_ent, f_ent.data(), sizeof(int32_t) * f_ent.size());
    if (!rc) rc = upload((void **)&d.j_pos, j_pos.data(), sizeof(int32_t) * j_pos.size());
    if (!rc) rc = upload((void **)&d.j_ptr, j_ptr.data(), sizeof(int32_t) * j_ptr.size());
    if (!rc) rc = upload((void **)&d.j_ent, j_ent.data(), sizeof(int32_t) * j_ent.size());
    if (!rc) rc = upload((void **)&d.r_ptr, r_ptr.data(), sizeof(int32_t) * r_ptr.size());
    if (!rc) rc = upload((void **)&d.r_z, r_z.data(), sizeof(int32_t) * r_z.size());
    if (!rc && !sc_ent.empty()) rc = upload((void **)&d.sc_ent, sc_ent.data(), sizeof(int32_t) * sc_ent.size());
    if (!rc && !sc_ent.empty()) rc = upload((void **)&d.sc_tile, sc_tile.data(), sizeof(uint32_t) * sc_tile.size());
    d.sc_len = sc_len;
    d.sc_cap = sc_cap;
    return rc;
}

void mass_action_release(wind_network *n) {
    MassActionDevice &d = n->ma;
    void *ptrs[] = {d.rx, d.k, d.f_ptr, d.f_ent, d.j_pos, d.j_ptr, d.j_ent, d.r_ptr, d.r_z, d.sc_ent, d.sc_tile};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    d = MassActionDevice();
}

// Which kernel launch_mass_action picks, and whether it needs the
// process-wide matrix scratch (only callers holding the context lock may
// use that).
struct MassActionPlan {
    bool dense = false;      // ma_dense_integrate
    bool in_lds = true;      // ma_block_integrate: matrices in LDS
    size_t lds = 0;
};

// PIVOT (fast mode, WIND_FLAG_PIVOT) takes the LDS LU with partial pivoting
// instead of the dense solver, whose diagonal blocks are unpivoted.
static MassActionPlan mass_action_plan(const MassActionDevice &d, int mode, bool pivot = false) {
    MassActionPlan p;
    int dev = 0, max_lds = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const int N = d.N;
    const bool fast = mode == WIND_MODE_FAST;
    if (fast && !pivot && (N + 15) / 16 * 16 <= 192) {
        const size_t lds = dense_lds_bytes(N, d.R);
        if (lds + 128 <= (size_t)max_lds) {
            p.dense = true;
            p.lds = lds;
            return p;
        }
    }
    const size_t base = sizeof(float) * (5 * (size_t)N + 3 * (size_t)d.R);
    const size_t mats = sizeof(float) * (fast ? 1 : 2) * (size_t)N * N;
    p.in_lds = base + mats + 64 <= (size_t)max_lds;
    p.lds = base + (p.in_lds ? mats : 0);
    return p;
}

bool mass_action_needs_scratch(const wind_network *net, int mode) {
    const MassActionPlan p = mass_action_plan(net->ma, mode);
    return !p.dense && !p.in_lds;
}

namespace {

// Launch shape of the one-workgroup-per-system kernel (ma_block_integrate).
struct BlockShape {
    int T = 0, grid = 0;
    size_t lds = 0;
    bool in_lds = true;
};

template <int METHOD>
auto block_kernel(bool fast, bool pivot) {
    return fast ? (pivot ? ma_block_integrate<METHOD, true, true> : ma_block_integrate<METHOD, true>)
                : ma_block_integrate<METHOD, false>;
}

template <int METHOD>
int block_shape(int num_cu, const MassActionDevice &d, const MassActionPlan &plan, bool fast, const SolverChoice &sc,
                int nsys, BlockShape *s) {
    s->T = block_threads(d.N, nsys, num_cu);
    s->in_lds = plan.in_lds;
    s->lds = plan.lds;
    int max_lds = 0, dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (s->lds + 64 > (size_t)max_lds) return wind_fail(WIND_ERR_UNSUPPORTED, "network too large for LDS");
    auto kfn = block_kernel<METHOD>(fast, sc.pivot != 0);
    if (s->lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)s->lds);
        if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, s->T, s->lds);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    s->grid = scaled_grid(num_cu, per_cu, sc, nsys);
    return WIND_OK;
}

size_t scratch_bytes(const MassActionDevice &d, const BlockShape &s) {
    return s.in_lds ? 0 : sizeof(float) * 2 * (size_t)d.N * d.N * s.grid;
}

}  // namespace

template <int METHOD>
int mass_action_scratch_bytes(int num_cu, const wind_network *net, int mode, const SolverChoice &sc, int nsys,
                              size_t *bytes) {
    *bytes = 0;
    if (net->builtin != WIND_NET_MASS_ACTION) return WIND_OK;
    const MassActionPlan plan = mass_action_plan(net->ma, mode, sc.pivot != 0);
    if (plan.dense || plan.in_lds) return WIND_OK;
    BlockShape s;
    const int rc = block_shape<METHOD>(num_cu, net->ma, plan, mode == WIND_MODE_FAST, sc, nsys, &s);
    if (rc) return rc;
    *bytes = scratch_bytes(net->ma, s);
    return WIND_OK;
}

// Grows the scratch; a growth frees the old buffer, so anything that kept
// its address (a captured batch graph) must key on mass_action_scratch_ptr().
int mass_action_scratch_reserve(size_t need) {
    if (need <= g_scratch.bytes) return WIND_OK;
    if (g_scratch.ptr) (void)hipFree(g_scratch.ptr);
    g_scratch.ptr = nullptr;
    g_scratch.bytes = 0;
    hipError_t e = hipMalloc((void **)&g_scratch.ptr, need);
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    g_scratch.bytes = need;
    return WIND_OK;
}

const void *mass_action_scratch_ptr() { return g_scratch.ptr; }

template <int METHOD>
int launch_mass_action(int num_cu, const wind_network *net, const IntegrateArgs &a, int mode,
                       int kernel, const SolverChoice &sc, hipStream_t st) {
    if (kernel == WIND_KERNEL_LANE)
        return wind_fail(WIND_ERR_UNSUPPORTED, "mass-action networks run on the block kernel");
    const MassActionDevice &d = net->ma;
    const int N = d.N;
    const bool fast = mode == WIND_MODE_FAST;
    const MassActionPlan plan = mass_action_plan(d, mode, sc.pivot != 0);
    if (plan.dense) {
        if (METHOD == WIND_SIE && dense_reg_eligible(d, sc)) return launch_dense_reg(num_cu, d, a, sc, st);
        const size_t lds = plan.lds;
        int T = ((N + 15) & ~15) <= 64 ? 64 : 512;
        if (sc.block == 64 || sc.block == 128 || sc.block == 256 || sc.block == 512) T = sc.block;
        const int solver = sc.solver == WIND_SOLVER_AUTO ? WIND_SOLVER_MFMA_F32 : sc.solver;
        switch (solver) {
        case WIND_SOLVER_VALU: return launch_dense<METHOD, PREC_VALU>(num_cu, d, a, sc.ir_sweeps, lds, T, sc, st);
        case WIND_SOLVER_MFMA_F32:
            return launch_dense<METHOD, PREC_MFMA_F32>(num_cu, d, a, sc.ir_sweeps, lds, T, sc, st);
        case WIND_SOLVER_MFMA_BF16:
            return launch_dense<METHOD, PREC_MFMA_BF16>(num_cu, d, a, sc.ir_sweeps, lds, T, sc, st);
        default:
            return launch_dense<METHOD, PREC_MFMA_BF16X3>(num_cu, d, a, sc.ir_sweeps, lds, T, sc, st);
        }
    }
    if (a.trace.capacity > 0)
        return wind_fail(WIND_ERR_UNSUPPORTED, "trace: mass-action networks are traced by the dense solver "
                                               "(fast mode, N <= 176) only");
    BlockShape s;
    int rc = block_shape<METHOD>(num_cu, d, plan, fast, sc, a.nsys, &s);
    if (rc) return rc;
    // Matrices beyond LDS live in the process-wide scratch, one slice per
    // workgroup.  Launches that use it must not overlap: callers hold the
    // context lock, and batches reserve it up front and put every bucket
    // that uses it on one stream.
    float *scratch = nullptr;
    if (!s.in_lds) {
        rc = mass_action_scratch_reserve(scratch_bytes(d, s));
        if (rc) return rc;
        scratch = g_scratch.ptr;
    }
    auto kfn = block_kernel<METHOD>(fast, sc.pivot != 0);
    hipLaunchKernelGGL(kfn, dim3(s.grid), dim3(s.T), s.lds, st, a, d, scratch, s.in_lds ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return wind_fail(WIND_ERR_DEVICE, hipGetErrorString(e));
    return WIND_OK;
}

template int launch_mass_action<WIND_SIE>(int, const wind_network *, const IntegrateArgs &, int, int,
                                          const SolverChoice &, hipStream_t);
template int launch_mass_action<WIND_RK2>(int, const wind_network *, const IntegrateArgs &, int, int,
                                          const SolverChoice &, hipStream_t);
template int mass_action_scratch_bytes<WIND_SIE>(int, const wind_network *, int, const SolverChoice &, int, size_t *);
template int mass_action_scratch_bytes<WIND_RK2>(int, const wind_network *, int, const SolverChoice &, int, size_t *);
