// mps_batch.hip -- batched form: many independent solves in one device arena, stepped
// together by persistent kernels replayed from a captured hipGraph.
//
// Layout: every solve ("instance") owns one contiguous block of the arena holding its SoA
// arrays (same layout as the single-solve path) and one MpsCtl in a separate control array.
// Work is described by two device tables of MpsWork items, never spanning two instances:
// element items are patches (MPS_PATCH elements, processed by one workgroup of elem_block
// work-items), node items are node ranges one workgroup wide.  The graph replays
//   begin, (element, node+begin)x(S-1), element, node
// over the tables; instances at different step counts coexist because each carries its own
// step protocol (MpsCtl), and finished instances are skipped by the go flags.
//
// Tables are double-buffered: an update goes to the buffer no queued launch reads, and the
// graph captured for that buffer is used from then on, so admitting or retiring instances
// never re-captures a graph and never stalls the compute stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mps_gpu.h"
#include "mps_api_internal.h"
#include "mps_kernels.h"

using namespace mpsh;

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, size) with coalescing of freed ranges.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes, noff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto& kv : free_) s += kv.second;
        return s;
    }
    size_t largest() const {
        size_t s = 0;
        for (auto& kv : free_) s = std::max(s, kv.second);
        return s;
    }
    size_t high_water() const {  // end of the last allocated byte
        size_t hw = size;
        auto it = free_.rbegin();
        if (it != free_.rend() && it->first + it->second == size) hw = it->first;
        return hw;
    }
};

struct Slot {
    bool used = false;
    size_t off = 0, bytes = 0;
    MpsInst inst{};
    int64_t n_steps = 0;
};

struct Table {
    MpsWork* elem = nullptr;  // device
    MpsWork* node = nullptr;
    int* count = nullptr;     // device: [0] element items, [1] node items
    int cap = 0;
    hipEvent_t last_use = nullptr;  // recorded after the last launch that read this buffer
    hipGraphExec_t exec = nullptr;  // graph captured for this buffer
};

}  // namespace

// Bytes of device state of one solve (arena block).
size_t mps_instance_bytes(const HostPrep& hp, int precision) {
    (void)precision;  // element forces are summed into FP64 partials in every precision form
    const size_t N = hp.N, Ne = hp.Ne, nen = hp.nen, NP = hp.n_part, NPA = hp.n_patch;
    const size_t nrec = hp.bcval.size() / 3;
    size_t b = 0;
    b += align_up(3 * N * 8) * 4;                 // X, u, v, a
    b += align_up(N * 8);                         // mass
    b += align_up(N * 4);                         // ninfo
    b += align_up((N + 1) * 4) * 2;               // rowptr, nrow
    b += align_up(Ne * nen * 4) * 2;              // conn, col
    b += align_up(Ne * 8);                        // l0
    b += align_up(std::max<size_t>(NP * 3 * 8, Ne * nen * 8));  // partial forces (also set-up scratch)
    b += align_up((NPA + 1) * 4);                 // poff
    b += align_up((NP + NPA) * 2);                // prow
    b += align_up(Ne * nen * 2);                  // pcol
    b += align_up(NP * 4);                        // ncol
    b += align_up(std::max<size_t>(nrec, 1) * 3 * 8);  // bcval
    b += align_up(2 * MPS_DT_SLOTS * MPS_DT_STRIDE * 8);  // dt reduction slots
    if (hp.per_elem) b += align_up(2 * Ne * 8) + align_up(Ne * 8) * 2;  // mat, cel, rho
    return b;
}

struct mps_batch_s {
    mps_batch_cfg cfg{};
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;  // compute
    hipStream_t up = nullptr;      // uploads (non-blocking)
    char* arena = nullptr;
    Arena alloc;
    MpsCtl* ctl = nullptr;         // device [max_instances]
    MpsInst* insts = nullptr;      // device [max_instances]
    std::vector<MpsCtl> ctl_h;     // mirror refreshed at sync points
    std::vector<Slot> slots;
    Table tab[2];
    int cur = 0;                   // table in use
    bool dirty = false;
    int nen = -1, law = -1;
    int bs = 0, grid_mult = 0;
    int steps_per_launch = 32;
    hipEvent_t last_launch = nullptr;
    bool any_launch = false;
    hipEvent_t tl0 = nullptr, tl1 = nullptr;  // around the graph launches of the last enqueue
    bool timed = false;
    std::string err;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

#define BCHK(b, call)                                                                                      \
    do {                                                                                                   \
        hipError_t _e = (call);                                                                            \
        if (_e != hipSuccess) return (b)->fail(MPS_ERR_HIP, fmt("%s: %s", #call, hipGetErrorString(_e))); \
    } while (0)

static void batch_release(mps_batch_s* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& t : b->tab) {
        if (t.exec) (void)hipGraphExecDestroy(t.exec);
        if (t.elem) (void)hipFree(t.elem);
        if (t.node) (void)hipFree(t.node);
        if (t.count) (void)hipFree(t.count);
        if (t.last_use) (void)hipEventDestroy(t.last_use);
    }
    if (b->arena) (void)hipFree(b->arena);
    if (b->ctl) (void)hipFree(b->ctl);
    if (b->insts) (void)hipFree(b->insts);
    if (b->last_launch) (void)hipEventDestroy(b->last_launch);
    if (b->tl0) (void)hipEventDestroy(b->tl0);
    if (b->tl1) (void)hipEventDestroy(b->tl1);
    if (b->up) (void)hipStreamDestroy(b->up);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

extern "C" int mps_batch_create(const mps_batch_cfg* cfg, mps_batch* out) {
    if (!cfg || !out) return MPS_ERR_ARG;
    *out = nullptr;
    if (cfg->precision != MPS_PREC_FP64 && cfg->precision != MPS_PREC_FP32_ELEM) return MPS_ERR_ARG;
    if (cfg->elem_block && cfg->elem_block != 64 && cfg->elem_block != 128 && cfg->elem_block != 256)
        return MPS_ERR_ARG;
    if (cfg->grid_mult < 0 || cfg->grid_mult > 64 || cfg->steps_per_launch < 0 || cfg->max_instances < 0)
        return MPS_ERR_ARG;
    auto* b = new mps_batch_s();
    b->cfg = *cfg;
    if (b->cfg.arena_bytes <= 0) b->cfg.arena_bytes = (int64_t)1 << 30;
    if (b->cfg.max_instances <= 0) b->cfg.max_instances = 1024;
    b->device = cfg->device;
    b->steps_per_launch = cfg->steps_per_launch > 0 ? cfg->steps_per_launch : 32;
    b->bs = cfg->elem_block > 0 ? cfg->elem_block : 256;
    b->grid_mult = cfg->grid_mult > 0 ? cfg->grid_mult : 2;
    int rc = MPS_OK;
    hipDeviceProp_t p;
    if (hipSetDevice(b->device) != hipSuccess || hipGetDeviceProperties(&p, b->device) != hipSuccess) {
        rc = MPS_ERR_HIP;
    } else {
        b->cus = p.multiProcessorCount;
        const int M = b->cfg.max_instances;
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&b->up, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->last_launch, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&b->tl0) != hipSuccess || hipEventCreate(&b->tl1) != hipSuccess ||
            hipMalloc(&b->arena, (size_t)b->cfg.arena_bytes) != hipSuccess ||
            hipMalloc(&b->ctl, sizeof(MpsCtl) * M) != hipSuccess ||
            hipMalloc(&b->insts, sizeof(MpsInst) * M) != hipSuccess) {
            rc = MPS_ERR_HIP;
        } else {
            b->alloc.init((size_t)b->cfg.arena_bytes);
            b->slots.resize(M);
            b->ctl_h.resize(M);
            for (auto& t : b->tab)
                if (hipEventCreateWithFlags(&t.last_use, hipEventDisableTiming) != hipSuccess) rc = MPS_ERR_HIP;
        }
    }
    if (rc != MPS_OK) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return MPS_OK;
}

// Wait until no queued launch can still read slot / arena data that is about to change.
static int wait_launches(mps_batch_s* b) {
    if (b->any_launch) BCHK(b, hipEventSynchronize(b->last_launch));
    return MPS_OK;
}

// Checks that a prepared solve can join batch b (fixed steps, same precision / element / law).
int mps_batch_check(mps_batch b, const HostPrep& hp, std::string* why) {
    if (hp.cfg.n_steps <= 0) {
        *why = "batched solves need a fixed step count (n_steps > 0)";
        return MPS_ERR_ARG;
    }
    if (hp.cfg.precision != b->cfg.precision) {
        *why = "precision differs from the batch";
        return MPS_ERR_ARG;
    }
    if (hp.cfg.energy_every != 0) {
        *why = "energy history is not kept in the batched form";
        return MPS_ERR_ARG;
    }
    if (b->nen >= 0 && (hp.nen != b->nen || hp.law != b->law)) {
        *why = "a batch holds one element type and one constitutive law";
        return MPS_ERR_ARG;
    }
    return MPS_OK;
}

size_t mps_batch_footprint(mps_batch b, const HostPrep& hp) {
    return mps_instance_bytes(hp, b->cfg.precision);
}

// Free slot and largest free arena block (admission by footprint).
void mps_batch_room(mps_batch b, int* free_slots, size_t* largest) {
    int n = 0;
    for (auto& s : b->slots) n += !s.used;
    *free_slots = n;
    *largest = b->alloc.largest();
}

extern "C" int mps_batch_add(mps_batch b, const mps_mesh* mesh, const mps_material* mat, const mps_bcs* bcs,
                             const mps_solver_cfg* cfg, int32_t* slot_out) {
    if (!b || !slot_out) return MPS_ERR_ARG;
    HostPrep hp;
    std::string e;
    int rc = prepare(mesh, mat, bcs, cfg, &hp, &e);
    if (rc) return b->fail(rc, e);
    return mps_batch_add_prepared(b, hp, slot_out);
}

// Admission of an already prepared solve: arena block, uploads, set-up kernels.
int mps_batch_add_prepared(mps_batch b, const HostPrep& hp, int32_t* slot_out) {
    const mps_solver_cfg* cfg = &hp.cfg;
    std::string e;
    int rc = mps_batch_check(b, hp, &e);
    if (rc) return b->fail(rc, e);
    if (b->nen >= 0 && (hp.nen != b->nen || hp.law != b->law))
        return b->fail(MPS_ERR_ARG, "a batch holds one element type and one constitutive law");
    int slot = -1;
    for (int s = 0; s < (int)b->slots.size(); s++)
        if (!b->slots[s].used) {
            slot = s;
            break;
        }
    if (slot < 0) return b->fail(MPS_ERR_CAPACITY, "no free instance slot");
    const size_t N = hp.N, Ne = hp.Ne, nen = hp.nen;
    const size_t bytes = mps_instance_bytes(hp, cfg->precision);
    const size_t NP = hp.n_part, NPA = hp.n_patch;
    size_t off = 0;
    if (!b->alloc.alloc(bytes, &off)) return b->fail(MPS_ERR_CAPACITY, fmt("arena full (%zu bytes requested)", bytes));
    (void)hipSetDevice(b->device);
    char* base = b->arena + off;
    size_t cursor = 0;
    auto take = [&](size_t nb) {
        char* p = base + cursor;
        cursor += align_up(nb);
        return p;
    };
    MpsInst I{};
    I.N = hp.N;
    I.Ne = hp.Ne;
    I.nen = hp.nen;
    I.law = hp.law;
    I.precision = cfg->precision;
    I.bc_mode = hp.bc_mode;
    I.vel_ramp = hp.vel_ramp;
    I.time_mode = MPS_TIME_FIXED;
    I.n_steps = cfg->n_steps;
    I.t0 = cfg->t0;
    I.tf = cfg->tf;
    I.scale = cfg->scale;
    I.mat_a = hp.pa;
    I.mat_b = hp.pb;
    I.c0 = hp.c0;
    I.rho0 = hp.rho0;
    I.dof1_node = (int32_t)(1 % N);
    I.dof1_dir = (int32_t)(1 / N);
    I.X = (double*)take(3 * N * 8);
    I.u = (double*)take(3 * N * 8);
    I.v = (double*)take(3 * N * 8);
    I.a = (double*)take(3 * N * 8);
    I.mass = (double*)take(N * 8);
    I.ninfo = (uint32_t*)take(N * 4);
    I.rowptr = (int32_t*)take((N + 1) * 4);
    I.conn = (int32_t*)take(Ne * nen * 4);
    I.col = (int32_t*)take(Ne * nen * 4);
    I.l0 = (double*)take(Ne * 8);
    I.part = (double*)take(std::max<size_t>(NP * 3 * 8, Ne * nen * 8));
    I.poff = (int32_t*)take((NPA + 1) * 4);
    I.prow = (uint16_t*)take((NP + NPA) * 2);
    I.pcol = (uint16_t*)take(Ne * nen * 2);
    I.nrow = (int32_t*)take((N + 1) * 4);
    I.ncol = (int32_t*)take(NP * 4);
    I.n_patch = hp.n_patch;
    I.n_part = hp.n_part;
    I.dtred = (unsigned long long*)take(2 * MPS_DT_SLOTS * MPS_DT_STRIDE * 8);
    I.bcval = (double*)take(hp.bcval.size() * 8);
    if (hp.per_elem) {
        I.mat = (double*)take(2 * Ne * 8);
        I.cel = (double*)take(Ne * 8);
        I.rho = (double*)take(Ne * 8);
    }
    I.ctl = b->ctl + slot;
    I.hist = nullptr;
    I.hist_cap = 0;
    if (cursor > bytes) {
        b->alloc.release(off, bytes);
        return b->fail(MPS_ERR_STATE, "internal: instance layout exceeds its block");
    }
    // uploads on the non-blocking upload stream; the host waits only for its own copies
    auto up = [&](void* dst, const void* src, size_t nb) {
        return hipMemcpyAsync(dst, src, nb, hipMemcpyHostToDevice, b->up);
    };
    hipError_t he = hipSuccess;
    if (he == hipSuccess) he = up((void*)I.X, hp.X.data(), 3 * N * 8);
    if (he == hipSuccess) he = up((void*)I.ninfo, hp.ninfo.data(), N * 4);
    if (he == hipSuccess) he = up((void*)I.rowptr, hp.rowptr.data(), (N + 1) * 4);
    if (he == hipSuccess) he = up((void*)I.conn, hp.conn.data(), Ne * nen * 4);
    if (he == hipSuccess) he = up((void*)I.col, hp.col.data(), Ne * nen * 4);
    if (he == hipSuccess) he = up((void*)I.poff, hp.poff.data(), (NPA + 1) * 4);
    if (he == hipSuccess) he = up((void*)I.prow, hp.prow.data(), (NP + NPA) * 2);
    if (he == hipSuccess) he = up((void*)I.pcol, hp.pcol.data(), Ne * nen * 2);
    if (he == hipSuccess) he = up((void*)I.nrow, hp.nrow.data(), (N + 1) * 4);
    if (he == hipSuccess) he = up((void*)I.ncol, hp.ncol.data(), NP * 4);
    if (he == hipSuccess) he = up((void*)I.bcval, hp.bcval.data(), hp.bcval.size() * 8);
    if (hp.per_elem) {
        if (he == hipSuccess) he = up((void*)I.mat, hp.mat.data(), 2 * Ne * 8);
        if (he == hipSuccess) he = up((void*)I.cel, hp.cel.data(), Ne * 8);
        if (he == hipSuccess) he = up((void*)I.rho, hp.rho.data(), Ne * 8);
    }
    if (he == hipSuccess) he = up(b->insts + slot, &I, sizeof(I));
    if (he == hipSuccess) he = hipMemsetAsync(I.u, 0, 3 * align_up(3 * N * 8), b->up);  // u, v, a: adjacent blocks
    if (he == hipSuccess) he = hipStreamSynchronize(b->up);
    // set-up kernels on the compute stream, after whatever is queued there
    unsigned long long* dt0 = &b->ctl[slot].dt0min;
    if (he == hipSuccess) he = mpsk::launch_fill_u64(dt0, 1, 0x7FF0000000000000ull, b->stream);
    if (he == hipSuccess) he = mpsk::launch_fill_u64(I.dtred, 2 * MPS_DT_SLOTS * MPS_DT_STRIDE, 0x7FF0000000000000ull, b->stream);
    if (he == hipSuccess) he = mpsk::launch_setup(I, I.part, (double*)I.l0, dt0, b->stream);
    if (he == hipSuccess) he = mpsk::launch_init_ctl(I, dt0, b->stream);
    // dt of step 1 from u0 = 0 (solverExplicit.cpp:176-178)
    if (he == hipSuccess) he = mpsk::launch_dt(I, I.t0 < 1E-12 ? I.u : nullptr, I.dtred + MPS_DT_SLOTS * MPS_DT_STRIDE, b->stream);
    if (he != hipSuccess) {
        b->alloc.release(off, bytes);
        return b->fail(MPS_ERR_HIP, fmt("admission: %s", hipGetErrorString(he)));
    }
    Slot& S = b->slots[slot];
    S.used = true;
    S.off = off;
    S.bytes = bytes;
    S.inst = I;
    S.n_steps = cfg->n_steps;
    memset(&b->ctl_h[slot], 0, sizeof(MpsCtl));
    b->ctl_h[slot].status = MPS_RUNNING;
    b->ctl_h[slot].A = 1;
    b->ctl_h[slot].nan_step = LLONG_MAX;
    b->nen = hp.nen;
    b->law = hp.law;
    b->dirty = true;
    *slot_out = slot;
    return MPS_OK;
}

// (Re)capture the step graph of table t for the current block size / grid / S.
static int capture(mps_batch_s* b, int t) {
    Table& T = b->tab[t];
    if (T.exec) {
        (void)hipGraphExecDestroy(T.exec);
        T.exec = nullptr;
    }
    const int grid = b->grid_mult * b->cus;
    const int S = b->steps_per_launch;
    const int prec = b->cfg.precision;
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = mpsk::launch_batch_phase(2, b->bs, prec, b->nen, b->law, b->insts, T.node, T.count + 1, grid, 0,
                                            0u, b->stream);
    for (int s = 0; s < S && e == hipSuccess; s++) {
        e = mpsk::launch_batch_phase(0, b->bs, prec, b->nen, b->law, b->insts, T.elem, T.count, grid, 0, 0u,
                                     b->stream);
        // the first node kernel finds v and a (the begin phase leaves them in place), the others the half-step
        // velocity; the last keeps v and a (no prediction)
        if (e == hipSuccess)
            e = mpsk::launch_batch_phase(1, b->bs, prec, b->nen, b->law, b->insts, T.node, T.count + 1, grid,
                                         s < S - 1 ? 1 : 0, s == 0 ? MPS_N_IN_VA : 0u, b->stream);
    }
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return b->fail(MPS_ERR_HIP, fmt("graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2)));
    }
    hipError_t e3 = hipGraphInstantiate(&T.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) return b->fail(MPS_ERR_HIP, fmt("graph instantiate: %s", hipGetErrorString(e3)));
    return MPS_OK;
}

// Rebuild the work tables into the buffer that no queued launch reads, and switch to it.
static int flush_tables(mps_batch_s* b) {
    if (!b->dirty) return MPS_OK;
    std::vector<MpsWork> we, wn;
    const int bs = b->bs;
    for (int s = 0; s < (int)b->slots.size(); s++) {
        const Slot& S = b->slots[s];
        if (!S.used) continue;
        for (int e0 = 0; e0 < S.inst.Ne; e0 += MPS_PATCH)  // one item per patch
            we.push_back({s, e0, std::min(e0 + MPS_PATCH, S.inst.Ne), e0 == 0});
        for (int n0 = 0; n0 < S.inst.N; n0 += bs) wn.push_back({s, n0, std::min(n0 + bs, S.inst.N), n0 == 0});
    }
    const int nt = 1 - b->cur;
    Table& T = b->tab[nt];
    BCHK(b, hipEventSynchronize(T.last_use));  // launches that read this buffer are done
    const int need = (int)std::max(we.size(), wn.size());
    bool recap = T.exec == nullptr;
    if (need > T.cap) {
        if (T.elem) (void)hipFree(T.elem);
        if (T.node) (void)hipFree(T.node);
        if (T.count) (void)hipFree(T.count);
        T.elem = T.node = nullptr;
        T.count = nullptr;
        int cap = std::max(1024, need + need / 2);
        BCHK(b, hipMalloc(&T.elem, sizeof(MpsWork) * cap));
        BCHK(b, hipMalloc(&T.node, sizeof(MpsWork) * cap));
        BCHK(b, hipMalloc(&T.count, sizeof(int) * 2));
        T.cap = cap;
        recap = true;
    }
    int counts[2] = {(int)we.size(), (int)wn.size()};
    if (!we.empty()) BCHK(b, hipMemcpyAsync(T.elem, we.data(), we.size() * sizeof(MpsWork), hipMemcpyHostToDevice, b->up));
    if (!wn.empty()) BCHK(b, hipMemcpyAsync(T.node, wn.data(), wn.size() * sizeof(MpsWork), hipMemcpyHostToDevice, b->up));
    BCHK(b, hipMemcpyAsync(T.count, counts, sizeof(counts), hipMemcpyHostToDevice, b->up));
    BCHK(b, hipStreamSynchronize(b->up));
    if (recap) {
        int rc = capture(b, nt);
        if (rc) return rc;
    }
    b->cur = nt;
    b->dirty = false;
    return MPS_OK;
}

// Enqueue `launches` graph launches (S steps each) on the compute stream.
static int enqueue(mps_batch_s* b, int64_t launches) {
    int rc = flush_tables(b);
    if (rc) return rc;
    Table& T = b->tab[b->cur];
    if (!T.exec) {
        rc = capture(b, b->cur);
        if (rc) return rc;
    }
    BCHK(b, hipEventRecord(b->tl0, b->stream));
    for (int64_t i = 0; i < launches; i++) BCHK(b, hipGraphLaunch(T.exec, b->stream));
    BCHK(b, hipEventRecord(b->tl1, b->stream));
    b->timed = launches > 0;
    BCHK(b, hipEventRecord(T.last_use, b->stream));
    BCHK(b, hipEventRecord(b->last_launch, b->stream));
    b->any_launch = true;
    return MPS_OK;
}

static int refresh(mps_batch_s* b) {
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(b->ctl_h.data(), b->ctl, sizeof(MpsCtl) * b->ctl_h.size(), hipMemcpyDeviceToHost));
    return MPS_OK;
}

extern "C" int mps_batch_step(mps_batch b, int64_t max_steps) {
    if (!b || max_steps < 0) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    const int64_t launches = (max_steps + b->steps_per_launch - 1) / b->steps_per_launch;
    int rc = enqueue(b, launches);
    if (rc) return rc;
    return refresh(b);
}

extern "C" int mps_batch_launch(mps_batch b, int64_t launches) {
    if (!b || launches < 0) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    return enqueue(b, launches);
}

extern "C" int mps_batch_sync(mps_batch b) {
    if (!b) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    return refresh(b);
}

extern "C" int mps_batch_query(mps_batch b, int32_t slot, mps_report* rep) {
    if (!b || !rep || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return MPS_ERR_ARG;
    const MpsCtl& c = b->ctl_h[slot];
    const Slot& S = b->slots[slot];
    long long done = c.A - 1;
    int status = c.status;
    if (c.nan_step != LLONG_MAX) {
        status = MPS_ERR_NAN;
        done = c.nan_step;
    } else if (status == MPS_RUNNING && done >= S.n_steps) {
        status = MPS_DONE;
    }
    rep->steps = done < 0 ? 0 : done;
    rep->time = done > 0 ? c.slot[done & 1].time : S.inst.t0;
    rep->dt = done > 0 ? c.slot[done & 1].dt : c.slot[0].dt;
    rep->status = status;
    rep->nonfinite = c.nonfinite;
    rep->device_ms = 0.0;
    return MPS_OK;
}

extern "C" int mps_batch_get_state(mps_batch b, int32_t slot, double* u, double* v, double* a) {
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    BCHK(b, hipStreamSynchronize(b->stream));
    const MpsInst& I = b->slots[slot].inst;
    const size_t bytes = 3 * (size_t)I.N * 8;
    if (u) BCHK(b, hipMemcpy(u, I.u, bytes, hipMemcpyDeviceToHost));
    if (v) BCHK(b, hipMemcpy(v, I.v, bytes, hipMemcpyDeviceToHost));
    if (a) BCHK(b, hipMemcpy(a, I.a, bytes, hipMemcpyDeviceToHost));
    return MPS_OK;
}

extern "C" int mps_batch_remove(mps_batch b, int32_t slot) {
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    int rc = wait_launches(b);  // queued launches may still read this slot's data
    if (rc) return rc;
    Slot& S = b->slots[slot];
    b->alloc.release(S.off, S.bytes);
    S = Slot();
    b->dirty = true;
    return MPS_OK;
}

extern "C" int mps_batch_info(mps_batch b, int64_t* free_bytes, int64_t* largest_block, int32_t* active) {
    if (!b) return MPS_ERR_ARG;
    if (free_bytes) *free_bytes = (int64_t)b->alloc.free_bytes();
    if (largest_block) *largest_block = (int64_t)b->alloc.largest();
    if (active) {
        int n = 0;
        for (auto& s : b->slots) n += s.used;
        *active = n;
    }
    return MPS_OK;
}

// Element-steps the next `launches` graph launches will perform: every active solve advances by
// steps_per_launch steps per launch until its own step count (progress from the last sync; a solve admitted
// since then starts at step 1).
double mps_batch_launch_work(mps_batch b, int64_t launches) {
    double w = 0.0;
    const int64_t S = (int64_t)b->steps_per_launch * launches;
    for (size_t s = 0; s < b->slots.size(); s++) {
        const Slot& sl = b->slots[s];
        if (!sl.used) continue;
        const MpsCtl& c = b->ctl_h[s];
        if (c.status != MPS_RUNNING || c.nan_step != LLONG_MAX) continue;
        const int64_t left = sl.n_steps - (c.A - 1);
        if (left > 0) w += (double)sl.inst.Ne * (double)std::min(left, S);
    }
    return w;
}

// Active elements (the mix size the serving tuner keys on).
int64_t mps_batch_active_elements(mps_batch b) {
    int64_t n = 0;
    for (auto& sl : b->slots)
        if (sl.used) n += sl.inst.Ne;
    return n;
}

// Device milliseconds of the graph launches of the last enqueue (valid after a sync); < 0 if none.
double mps_batch_last_launch_ms(mps_batch b) {
    if (!b->timed) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->tl0, b->tl1) != hipSuccess) return -1.0;
    return ms;
}

void mps_batch_tuning(mps_batch b, int* elem_block, int* grid_mult) {
    *elem_block = b->bs;
    *grid_mult = b->grid_mult;
}

// Element block size and resident-block multiple for the following launches.  Results do not depend on
// them (every sum has a fixed order), so they may change between launches of a running batch.  Call only
// when no launch is in flight (after a sync): the captured graphs are released and re-captured.
int mps_batch_set_tuning(mps_batch b, int elem_block, int grid_mult) {
    if (elem_block != 64 && elem_block != 128 && elem_block != 256) return MPS_ERR_ARG;
    if (grid_mult < 1 || grid_mult > 64) return MPS_ERR_ARG;
    if (elem_block == b->bs && grid_mult == b->grid_mult) return MPS_OK;
    (void)hipSetDevice(b->device);
    int rc = wait_launches(b);
    if (rc) return rc;
    b->bs = elem_block;
    b->grid_mult = grid_mult;
    b->dirty = true;
    for (auto& t : b->tab)
        if (t.exec) {
            (void)hipGraphExecDestroy(t.exec);
            t.exec = nullptr;
        }
    return MPS_OK;
}

// Autotuner: times S-step graph launches for every (block size, grid multiple) on the current
// content, restores the state after each trial, keeps the fastest.
extern "C" int mps_batch_autotune(mps_batch b, int32_t* elem_block, int32_t* grid_mult) {
    if (!b) return MPS_ERR_ARG;
    (void)hipSetDevice(b->device);
    int rc = refresh(b);
    if (rc) return rc;
    const size_t used = b->alloc.high_water();
    char* backup = nullptr;
    MpsCtl* ctl_bk = nullptr;
    const size_t nctl = b->ctl_h.size();
    BCHK(b, hipMalloc(&backup, std::max<size_t>(used, 256)));
    if (hipMalloc(&ctl_bk, sizeof(MpsCtl) * nctl) != hipSuccess) {
        (void)hipFree(backup);
        return b->fail(MPS_ERR_HIP, "autotune backup allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto snapshot = [&](bool save) {
        if (save) {
            (void)hipMemcpyAsync(backup, b->arena, used, hipMemcpyDeviceToDevice, b->stream);
            (void)hipMemcpyAsync(ctl_bk, b->ctl, sizeof(MpsCtl) * nctl, hipMemcpyDeviceToDevice, b->stream);
        } else {
            (void)hipMemcpyAsync(b->arena, backup, used, hipMemcpyDeviceToDevice, b->stream);
            (void)hipMemcpyAsync(b->ctl, ctl_bk, sizeof(MpsCtl) * nctl, hipMemcpyDeviceToDevice, b->stream);
        }
    };
    snapshot(true);
    const int bs0 = b->bs, gm0 = b->grid_mult;
    float best = 1e30f;
    int best_bs = bs0, best_gm = gm0;
    const int bss[3] = {64, 128, 256};
    const int gms[5] = {1, 2, 4, 8, 16};
    for (int bs : bss)
        for (int gm : gms) {
            if (bs * gm > 4096) continue;  // more resident threads per CU than the hardware holds
            b->bs = bs;
            b->grid_mult = gm;
            b->dirty = true;
            for (auto& t : b->tab)
                if (t.exec) {
                    (void)hipGraphExecDestroy(t.exec);
                    t.exec = nullptr;
                }
            rc = flush_tables(b);
            if (rc == MPS_OK) rc = enqueue(b, 1);  // warm-up launch
            float ms = 0.f;
            if (rc == MPS_OK) {
                (void)hipEventRecord(e0, b->stream);
                rc = enqueue(b, 2);
                (void)hipEventRecord(e1, b->stream);
                (void)hipEventSynchronize(e1);
                (void)hipEventElapsedTime(&ms, e0, e1);
            }
            snapshot(false);
            if (rc != MPS_OK) break;
            if (ms < best) {
                best = ms;
                best_bs = bs;
                best_gm = gm;
            }
        }
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(backup);
    (void)hipFree(ctl_bk);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->bs = best_bs;
    b->grid_mult = best_gm;
    b->dirty = true;
    for (auto& t : b->tab)
        if (t.exec) {
            (void)hipGraphExecDestroy(t.exec);
            t.exec = nullptr;
        }
    if (rc) return rc;
    rc = flush_tables(b);
    if (rc) return rc;
    rc = refresh(b);
    if (elem_block) *elem_block = best_bs;
    if (grid_mult) *grid_mult = best_gm;
    return rc;
}

extern "C" int mps_batch_destroy(mps_batch b) {
    if (!b) return MPS_ERR_ARG;
    batch_release(b);
    delete b;
    return MPS_OK;
}

extern "C" const char* mps_batch_last_error(mps_batch b) { return b ? b->err.c_str() : ""; }
