// mps_api.hip -- C-ABI of libmps_gpu: single-solve path (see include/mps_gpu.h).
//
// Host responsibilities: validate inputs, build the node->(element,slot) CSR, the
// Dirichlet records and per-element material constants, upload, run the set-up
// kernels, then drive the step kernels (optionally through a captured hipGraph)
// without any per-step host round trip.  Every failure returns a status code;
// nothing here calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mps_gpu.h"
#include "mps_api_internal.h"
#include "mps_kernels.h"

namespace {

std::mutex g_err_mu;
std::string g_err;  // last mps_create failure

void set_global_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = buf;
}

}  // namespace

// ---------------------------------------------------------------------------
// Shared host preparation (also used by the batched form)
// ---------------------------------------------------------------------------
namespace mpsh {

std::string fmt(const char* f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

// Material constants exactly as the reference constructors compute them
// (hyperElasticNeoHookean.cpp:28-30, 180-181; soundSpeed :51-54).
void material_consts(int law, double E, double nu, double rho, double* pa, double* pb, double* c) {
    double lambda = (E * nu) / ((1. + nu) * (1. - 2. * nu));
    double mu = 0.5 * E / (1. + nu);
    double K1 = E / (3 * (1 - 2 * nu));
    *c = std::sqrt((lambda + 2 * mu) / rho);
    if (law == MPS_LAW_SVK || law == MPS_LAW_LINEAR_ELASTIC) {
        *c = std::sqrt(E / rho);  // hyperElasticStVenantKirchhoff.cpp soundSpeed, linearElastic.cpp:49-54
        *pa = lambda;
        *pb = mu;
    } else if (law == MPS_LAW_NH_COMPRESSIBLE) {
        *pa = lambda;
        *pb = mu;
    } else {
        *pa = mu;
        *pb = K1;
    }
}

int prepare(const mps_mesh* mesh, const mps_material* mat, const mps_bcs* bcs, const mps_solver_cfg* cfg,
            HostPrep* hp, std::string* err) {
    if (!mesh || !mat || !bcs || !cfg) {
        *err = "null argument";
        return MPS_ERR_ARG;
    }
    const int64_t N = mesh->n_nodes, Ne = mesh->n_elem, nen = mesh->nen;
    if (N <= 0 || Ne <= 0 || (nen != 8 && nen != 4) || !mesh->X || !mesh->conn) {
        *err = fmt("invalid mesh (n_nodes=%lld n_elem=%lld nen=%lld)", (long long)N, (long long)Ne, (long long)nen);
        return MPS_ERR_ARG;
    }
    if (N > (int64_t)(INT_MAX / 4) || Ne * nen > (int64_t)INT_MAX) {
        *err = "mesh too large for 32-bit indices";
        return MPS_ERR_ARG;
    }
    if (mat->law < MPS_LAW_NH_COMPRESSIBLE || mat->law > MPS_LAW_LINEAR_ELASTIC) {
        *err = fmt("unknown constitutive law %d", mat->law);
        return MPS_ERR_ARG;
    }
    if (!mat->E || !mat->nu || !mat->rho) {
        *err = "material arrays missing";
        return MPS_ERR_ARG;
    }
    if (cfg->precision != MPS_PREC_FP64 && cfg->precision != MPS_PREC_FP32_ELEM) {
        *err = fmt("unknown precision form %d", cfg->precision);
        return MPS_ERR_ARG;
    }
    if (!(cfg->scale > 0.0)) {
        *err = "scale factor must be > 0";
        return MPS_ERR_ARG;
    }
    if (cfg->n_steps < 0 || (cfg->n_steps == 0 && !(cfg->tf > cfg->t0))) {
        *err = "need n_steps > 0 or a time window tf > t0";
        return MPS_ERR_ARG;
    }
    if (bcs->mode < MPS_BC_RAMP || bcs->mode > MPS_BC_RAMP_RATE) {
        *err = fmt("unknown Dirichlet mode %d", bcs->mode);
        return MPS_ERR_ARG;
    }
    if (bcs->mode == MPS_BC_RAMP && bcs->n_dirichlet > 0 && !(cfg->tf - cfg->t0 != 0.0)) {
        *err = "ramp Dirichlet BCs need tf != t0 (UU = U/(tf-t0))";
        return MPS_ERR_ARG;
    }
    if (bcs->n_dirichlet < 0 || (bcs->n_dirichlet > 0 && (!bcs->dof || !bcs->value))) {
        *err = "invalid Dirichlet arrays";
        return MPS_ERR_ARG;
    }
    hp->N = (int32_t)N;
    hp->Ne = (int32_t)Ne;
    hp->nen = (int32_t)nen;
    hp->law = mat->law;
    // connectivity + CSR (ascending element, then slot)
    hp->conn.assign(mesh->conn, mesh->conn + Ne * nen);
    hp->rowptr.assign(N + 1, 0);
    for (int64_t k = 0; k < Ne * nen; k++) {
        int32_t n = hp->conn[k];
        if (n < 0 || n >= N) {
            *err = fmt("element %lld references node %d outside [0,%lld)", (long long)(k / nen), n, (long long)N);
            return MPS_ERR_ARG;
        }
        hp->rowptr[n + 1]++;
    }
    for (int64_t n = 0; n < N; n++) hp->rowptr[n + 1] += hp->rowptr[n];
    hp->col.resize(Ne * nen);
    {
        std::vector<int32_t> fill(hp->rowptr.begin(), hp->rowptr.end() - 1);
        for (int64_t k = 0; k < Ne * nen; k++) hp->col[fill[hp->conn[k]]++] = (int32_t)k;
    }
    // patches: MPS_PATCH consecutive elements each; distinct nodes ascending, their
    // contributions (local element, slot) ascending; node -> partials ascending by patch
    {
        const int64_t P = MPS_PATCH, np = (Ne + P - 1) / P;
        hp->n_patch = (int32_t)np;
        hp->poff.assign(np + 1, 0);
        hp->pcol.resize(Ne * nen);
        hp->prow.clear();
        hp->prow.reserve((size_t)(3 * Ne + np));
        std::vector<int32_t> pnode;
        pnode.reserve((size_t)(3 * Ne));
        std::vector<uint64_t> key((size_t)(P * nen));
        for (int64_t p = 0; p < np; p++) {
            const int64_t e0 = p * P, m = (std::min(e0 + P, Ne) - e0) * nen;
            for (int64_t q = 0; q < m; q++) key[q] = ((uint64_t)hp->conn[e0 * nen + q] << 16) | (uint64_t)q;
            std::sort(key.begin(), key.begin() + m);
            int64_t U = 0;
            for (int64_t q = 0; q < m; q++) {
                const int32_t n = (int32_t)(key[q] >> 16);
                if (q == 0 || n != pnode.back()) {
                    hp->prow.push_back((uint16_t)q);
                    pnode.push_back(n);
                    U++;
                }
                hp->pcol[e0 * nen + q] = (uint16_t)(key[q] & 0xffffu);
            }
            hp->prow.push_back((uint16_t)m);
            hp->poff[p + 1] = hp->poff[p] + (int32_t)U;
        }
        hp->n_part = hp->poff[np];
        hp->nrow.assign(N + 1, 0);
        for (int32_t n : pnode) hp->nrow[n + 1]++;
        for (int64_t n = 0; n < N; n++) hp->nrow[n + 1] += hp->nrow[n];
        hp->ncol.resize(pnode.size());
        std::vector<int32_t> fill(hp->nrow.begin(), hp->nrow.end() - 1);
        for (size_t g = 0; g < pnode.size(); g++) hp->ncol[fill[pnode[g]]++] = (int32_t)g;
    }
    // materials
    hp->per_elem = mat->per_element != 0;
    if (hp->per_elem) {
        hp->mat.resize(2 * Ne);
        hp->cel.resize(Ne);
        hp->rho.resize(Ne);
        for (int64_t e = 0; e < Ne; e++) {
            double E = mat->E[e], nu = mat->nu[e], rho = mat->rho[e];
            if (!(E > 0) || !(nu > -1.0 && nu < 0.5) || !(rho > 0)) {
                *err = fmt("element %lld: invalid material E=%g nu=%g rho=%g", (long long)e, E, nu, rho);
                return MPS_ERR_ARG;
            }
            material_consts(mat->law, E, nu, rho, &hp->mat[2 * e], &hp->mat[2 * e + 1], &hp->cel[e]);
            hp->rho[e] = rho;
        }
    } else {
        double E = mat->E[0], nu = mat->nu[0], rho = mat->rho[0];
        if (!(E > 0) || !(nu > -1.0 && nu < 0.5) || !(rho > 0)) {
            *err = fmt("invalid material E=%g nu=%g rho=%g", E, nu, rho);
            return MPS_ERR_ARG;
        }
        material_consts(mat->law, E, nu, rho, &hp->pa, &hp->pb, &hp->c0);
        hp->rho0 = rho;
    }
    // Dirichlet records: one per constrained node; value per constrained direction
    hp->ninfo.assign(N, 0u);
    std::vector<int32_t> rec_of(N, -1);
    hp->bcval.clear();
    int32_t nrec = 0;
    for (int32_t k = 0; k < bcs->n_dirichlet; k++) {
        int64_t dof = bcs->dof[k];
        if (dof < 0 || dof >= 3 * N) {
            *err = fmt("Dirichlet record %d: DOF %lld outside [0,%lld)", k, (long long)dof, (long long)(3 * N));
            return MPS_ERR_ARG;
        }
        int32_t n = (int32_t)(dof % N), d = (int32_t)(dof / N);
        if (hp->ninfo[n] & (1u << d)) {
            *err = fmt("duplicate Dirichlet DOF %lld (node %d dir %d)", (long long)dof, n + 1, d);
            return MPS_ERR_ARG;
        }
        if (rec_of[n] < 0) {
            rec_of[n] = nrec++;
            hp->bcval.resize(3 * (size_t)nrec, 0.0);
        }
        double val = bcs->value[k];
        if (bcs->mode == MPS_BC_RAMP) val = val / (cfg->tf - cfg->t0);  // UU = U/duration (classNodes.cpp:70)
        hp->bcval[3 * (size_t)rec_of[n] + d] = val;
        hp->ninfo[n] |= (1u << d);
    }
    if (nrec > (int32_t)(MPS_REC_MASK >> 3)) {
        *err = fmt("%d Dirichlet nodes: at most %u", nrec, MPS_REC_MASK >> 3);
        return MPS_ERR_ARG;
    }
    for (int64_t n = 0; n < N; n++)
        if (rec_of[n] >= 0) hp->ninfo[n] |= ((uint32_t)rec_of[n] << 3);
    if (hp->bcval.empty()) hp->bcval.assign(3, 0.0);
    hp->X.assign(mesh->X, mesh->X + 3 * N);
    hp->bc_mode = bcs->mode == MPS_BC_RAMP_RATE ? MPS_BC_RAMP : bcs->mode;  // a given rate runs as a ramp
    hp->vel_ramp = bcs->vel_ramp_steps;
    hp->cfg = *cfg;
    return MPS_OK;
}

}  // namespace mpsh

using namespace mpsh;

// ---------------------------------------------------------------------------
// Solver handle
// ---------------------------------------------------------------------------
struct mps_solver_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    MpsInst I{};
    mps_solver_cfg cfg{};
    std::vector<void*> allocs;
    double* scratch_u = nullptr;
    double* scratch_f = nullptr;
    unsigned long long* scratch_dt = nullptr;
    double* eout = nullptr;       // mps_element_outputs buffer (allocated on first use)
    int32_t* pick_idx = nullptr;  // mps_get_dofs buffers (grow-only, freed with the handle)
    double* pick_out = nullptr;
    int64_t pick_cap = 0;
    std::vector<void*> load_allocs;  // mps_set_loads buffers (replaced by the next call)
    int energy_nb = 0;
    bool primed = false;       // prologue (first dt) done
    long long next_step = 1;   // host mirror of ctl->A
    bool step_begun = false;   // step next_step already predicted on the device
    std::vector<std::pair<long long, hipGraphExec_t>> graphs;  // captured chunks by step count
    bool use_graphs = true;
    // optional kernel timing (HIP events around each element / node launch)
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    size_t ev_used = 0;
    double kt_ms[4] = {0, 0, 0, 0};   // element ms, node ms, begin ms, energy ms
    long long kt_n[4] = {0, 0, 0, 0};
    std::string err;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

#define HCHK(h, call)                                                                              \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) return (h)->fail(MPS_ERR_HIP, fmt("%s: %s", #call, hipGetErrorString(_e))); \
    } while (0)

template <typename P>
static hipError_t dalloc(mps_solver_s* h, P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes < 256 ? 256 : bytes);
    if (e == hipSuccess) {
        h->allocs.push_back(q);
        *p = (P*)q;
    }
    return e;
}

static void release(mps_solver_s* h) {
    if (!h) return;
    for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
    h->evpool.clear();
    if (h->device >= 0) (void)hipSetDevice(h->device);
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
    h->graphs.clear();
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    for (void* p : h->load_allocs) (void)hipFree(p);
    h->load_allocs.clear();
    if (h->pick_idx) (void)hipFree(h->pick_idx);
    if (h->pick_out) (void)hipFree(h->pick_out);
    if (h->eout) (void)hipFree(h->eout);
    h->pick_idx = nullptr;
    h->pick_out = nullptr;
    h->eout = nullptr;
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

static int upload_instance(mps_solver_s* h, const HostPrep& hp) {
    MpsInst& I = h->I;
    const size_t N = hp.N, Ne = hp.Ne, nen = hp.nen;
    I.N = hp.N;
    I.Ne = hp.Ne;
    I.nen = hp.nen;
    I.law = hp.law;
    I.precision = hp.cfg.precision;
    I.bc_mode = hp.bc_mode;
    I.vel_ramp = hp.vel_ramp;
    I.time_mode = hp.cfg.n_steps > 0 ? MPS_TIME_FIXED : MPS_TIME_WINDOW;
    I.n_steps = hp.cfg.n_steps;
    I.t0 = hp.cfg.t0;
    I.tf = hp.cfg.tf;
    I.scale = hp.cfg.scale;
    I.mat_a = hp.pa;
    I.mat_b = hp.pb;
    I.c0 = hp.c0;
    I.rho0 = hp.rho0;
    I.dof1_node = (int32_t)(1 % N);
    I.dof1_dir = (int32_t)(1 / N);
    double *X, *u, *v, *a, *mass, *l0, *matp = nullptr, *cel = nullptr, *rho = nullptr, *bcval, *hist, *part;
    double* pforce;
    int32_t *conn, *rowptr, *col, *poff, *nrow, *ncol;
    uint16_t *prow, *pcol;
    uint32_t* ninfo;
    MpsCtl* ctl;
    const size_t NP = (size_t)hp.n_part, NPA = (size_t)hp.n_patch;
    const size_t pf_bytes = std::max(NP * 3 * 8, Ne * nen * 8);  // partial forces; set-up scratch first
    const int cap = hp.cfg.history_cap > 0 ? hp.cfg.history_cap : 4096;
    HCHK(h, dalloc(h, &X, 3 * N * 8));
    HCHK(h, dalloc(h, &u, 3 * N * 8));
    HCHK(h, dalloc(h, &v, 3 * N * 8));
    HCHK(h, dalloc(h, &a, 3 * N * 8));
    HCHK(h, dalloc(h, &mass, N * 8));
    HCHK(h, dalloc(h, &l0, Ne * 8));
    HCHK(h, dalloc(h, &conn, Ne * nen * 4));
    HCHK(h, dalloc(h, &rowptr, (N + 1) * 4));
    HCHK(h, dalloc(h, &col, Ne * nen * 4));
    HCHK(h, dalloc(h, &ninfo, N * 4));
    HCHK(h, dalloc(h, &bcval, hp.bcval.size() * 8));
    HCHK(h, dalloc(h, &pforce, pf_bytes));
    HCHK(h, dalloc(h, &poff, (NPA + 1) * 4));
    HCHK(h, dalloc(h, &prow, (NP + NPA) * 2));
    HCHK(h, dalloc(h, &pcol, Ne * nen * 2));
    HCHK(h, dalloc(h, &nrow, (N + 1) * 4));
    HCHK(h, dalloc(h, &ncol, NP * 4));
    HCHK(h, dalloc(h, &ctl, sizeof(MpsCtl)));
    unsigned long long* dtred;
    HCHK(h, dalloc(h, &dtred, 2 * MPS_DT_SLOTS * MPS_DT_STRIDE * 8));
    HCHK(h, dalloc(h, &hist, 5 * (size_t)cap * 8));
    h->energy_nb = (int)((Ne + 255) / 256);
    HCHK(h, dalloc(h, &part, 3 * (size_t)h->energy_nb * 8));
    HCHK(h, dalloc(h, &h->scratch_u, 3 * N * 8));
    HCHK(h, dalloc(h, &h->scratch_f, 3 * N * 8));
    HCHK(h, dalloc(h, &h->scratch_dt, 64));
    if (hp.per_elem) {
        HCHK(h, dalloc(h, &matp, 2 * Ne * 8));
        HCHK(h, dalloc(h, &cel, Ne * 8));
        HCHK(h, dalloc(h, &rho, Ne * 8));
        HCHK(h, hipMemcpyAsync(matp, hp.mat.data(), 2 * Ne * 8, hipMemcpyHostToDevice, h->stream));
        HCHK(h, hipMemcpyAsync(cel, hp.cel.data(), Ne * 8, hipMemcpyHostToDevice, h->stream));
        HCHK(h, hipMemcpyAsync(rho, hp.rho.data(), Ne * 8, hipMemcpyHostToDevice, h->stream));
    }
    HCHK(h, hipMemcpyAsync(X, hp.X.data(), 3 * N * 8, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemsetAsync(u, 0, 3 * N * 8, h->stream));
    HCHK(h, hipMemsetAsync(v, 0, 3 * N * 8, h->stream));
    HCHK(h, hipMemsetAsync(a, 0, 3 * N * 8, h->stream));
    HCHK(h, hipMemcpyAsync(conn, hp.conn.data(), Ne * nen * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(rowptr, hp.rowptr.data(), (N + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(col, hp.col.data(), Ne * nen * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(poff, hp.poff.data(), (NPA + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(prow, hp.prow.data(), (NP + NPA) * 2, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(pcol, hp.pcol.data(), Ne * nen * 2, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(nrow, hp.nrow.data(), (N + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(ncol, hp.ncol.data(), NP * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(ninfo, hp.ninfo.data(), N * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(bcval, hp.bcval.data(), hp.bcval.size() * 8, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemsetAsync(hist, 0xFF, 5 * (size_t)cap * 8, h->stream));  // NaN
    I.X = X;
    I.conn = conn;
    I.l0 = l0;
    I.mat = matp;
    I.cel = cel;
    I.rho = rho;
    I.u = u;
    I.v = v;
    I.a = a;
    I.mass = mass;
    I.ninfo = ninfo;
    I.bcval = bcval;
    I.rowptr = rowptr;
    I.col = col;
    I.poff = poff;
    I.prow = prow;
    I.pcol = pcol;
    I.nrow = nrow;
    I.ncol = ncol;
    I.part = pforce;
    I.dtred = dtred;
    I.n_patch = hp.n_patch;
    I.n_part = hp.n_part;
    I.ctl = ctl;
    I.hist = hist;
    I.partials = part;
    I.hist_cap = cap;
    return MPS_OK;
}

// Set-up kernels: l0, lumped mass, dt0 = s*min(l0/c); initial control block.
static int device_setup(mps_solver_s* h) {
    MpsInst& I = h->I;
    double* mslot = I.part;  // per-(element, slot) mass values; the partial-force buffer is free until the first step
    unsigned long long inf = 0x7FF0000000000000ull;
    HCHK(h, hipMemcpyAsync(h->scratch_dt, &inf, 8, hipMemcpyHostToDevice, h->stream));
    hipError_t e = mpsk::launch_setup(I, mslot, (double*)I.l0, h->scratch_dt, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("set-up kernels: %s", hipGetErrorString(e)));
    unsigned long long bits = 0;
    HCHK(h, hipMemcpy(&bits, h->scratch_dt, 8, hipMemcpyDeviceToHost));
    double dt0;
    memcpy(&dt0, &bits, 8);
    MpsCtl c{};
    c.slot[0].time = I.t0;
    c.slot[0].dt = dt0 * I.scale;  // solverExplicit.cpp:165-166
    c.slot[0].go = 1;
    c.slot[1].go = 0;
    c.A = 1;
    c.B = 0;
    c.status = MPS_RUNNING;
    c.nonfinite = 0;
    c.nan_step = LLONG_MAX;
    c.dt0min = bits;
    HCHK(h, hipMemcpy(I.ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    hipError_t ef = mpsk::launch_fill_u64(I.dtred, 2 * MPS_DT_SLOTS * MPS_DT_STRIDE, inf, h->stream);
    if (ef == hipSuccess) ef = hipStreamSynchronize(h->stream);
    if (ef != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("dt slots: %s", hipGetErrorString(ef)));
    return MPS_OK;
}

extern "C" int mps_create(const mps_mesh* mesh, const mps_material* mat, const mps_bcs* bcs,
                          const mps_solver_cfg* cfg, mps_handle* out) {
    if (!out) {
        set_global_error("mps_create: out is NULL");
        return MPS_ERR_ARG;
    }
    *out = nullptr;
    HostPrep hp;
    std::string err;
    int rc = prepare(mesh, mat, bcs, cfg, &hp, &err);
    if (rc != MPS_OK) {
        set_global_error("mps_create: %s", err.c_str());
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_global_error("mps_create: no HIP device available");
        return MPS_ERR_HIP;
    }
    if (cfg->device < 0 || cfg->device >= ndev) {
        set_global_error("mps_create: device %d out of range (%d devices)", cfg->device, ndev);
        return MPS_ERR_ARG;
    }
    mps_solver_s* h = new mps_solver_s();
    h->device = cfg->device;
    h->cfg = *cfg;
    if (hipSetDevice(h->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        set_global_error("mps_create: HIP stream/event creation failed");
        release(h);
        delete h;
        return MPS_ERR_HIP;
    }
    rc = upload_instance(h, hp);
    if (rc == MPS_OK) rc = device_setup(h);
    if (rc != MPS_OK) {
        set_global_error("mps_create: %s", h->err.c_str());
        release(h);
        delete h;
        return rc;
    }
    *out = h;
    return MPS_OK;
}

extern "C" int mps_set_state(mps_handle h, const double* u, const double* v, const double* a) {
    if (!h) return MPS_ERR_ARG;
    if (h->primed || h->next_step != 1) return h->fail(MPS_ERR_STATE, "mps_set_state: only valid before the first step");
    (void)hipSetDevice(h->device);
    const size_t bytes = 3 * (size_t)h->I.N * 8;
    double* dst[3] = {h->I.u, h->I.v, h->I.a};
    const double* src[3] = {u, v, a};
    for (int i = 0; i < 3; i++) {
        if (src[i]) HCHK(h, hipMemcpyAsync(dst[i], src[i], bytes, hipMemcpyHostToDevice, h->stream));
        else HCHK(h, hipMemsetAsync(dst[i], 0, bytes, h->stream));
    }
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

// Nodal force entries are sorted by node (stable, so each node keeps its entries in input order) and
// flagged in ninfo; gravity is summed per node once, on the device.
extern "C" int mps_set_loads(mps_handle h, const mps_loads* L) {
    if (!h) return MPS_ERR_ARG;
    if (h->primed || h->next_step != 1) return h->fail(MPS_ERR_STATE, "mps_set_loads: only valid before the first step");
    const int64_t nf = L ? L->n_force : 0;
    const double grav = L ? L->gravity : 0.0;
    const int64_t N = h->I.N;
    if (nf < 0) return h->fail(MPS_ERR_ARG, "mps_set_loads: n_force < 0");
    if (!std::isfinite(grav)) return h->fail(MPS_ERR_ARG, "mps_set_loads: gravity is not finite");
    if (nf > 0 && (!L->node || !L->force || !L->t0 || !L->tf))
        return h->fail(MPS_ERR_ARG, "mps_set_loads: n_force > 0 needs node, force, t0 and tf");
    for (int64_t k = 0; k < nf; k++) {
        if (L->node[k] < 0 || L->node[k] >= N)
            return h->fail(MPS_ERR_ARG, fmt("mps_set_loads: force entry %lld: node %d outside [0,%lld)", (long long)k,
                                            L->node[k], (long long)N));
        if (!std::isfinite(L->t0[k]) || !std::isfinite(L->tf[k]) || !(L->tf[k] > L->t0[k]))
            return h->fail(MPS_ERR_ARG, fmt("mps_set_loads: force entry %lld: window needs finite tf > t0", (long long)k));
        for (int d = 0; d < 3; d++)
            if (!std::isfinite(L->force[3 * k + d]))
                return h->fail(MPS_ERR_ARG, fmt("mps_set_loads: force entry %lld is not finite", (long long)k));
    }
    (void)hipSetDevice(h->device);
    MpsInst& I = h->I;
    // host copy of the node flags: clear the loaded bit, set it on the nodes with entries
    std::vector<uint32_t> ninfo((size_t)N);
    HCHK(h, hipMemcpy(ninfo.data(), I.ninfo, (size_t)N * 4, hipMemcpyDeviceToHost));
    for (uint32_t& f : ninfo) f &= MPS_REC_MASK;
    std::vector<int32_t> order((size_t)nf);
    for (int64_t k = 0; k < nf; k++) order[(size_t)k] = (int32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return L->node[x] < L->node[y]; });
    std::vector<int32_t> lnode((size_t)nf), lrec((size_t)nf);
    std::vector<double> lm(3 * (size_t)nf), lwin(2 * (size_t)nf);
    for (size_t j = 0; j < (size_t)nf; j++) {
        const int32_t k = order[j];
        lnode[j] = L->node[k];
        lrec[j] = k;
        for (int d = 0; d < 3; d++) lm[3 * j + d] = L->force[3 * (size_t)k + d] / (L->tf[k] - L->t0[k]);  // classNodes.cpp:93-95
        lwin[2 * j] = L->t0[k];
        lwin[2 * j + 1] = L->tf[k];
        ninfo[(size_t)lnode[j]] |= MPS_NODE_LOADED;
    }
    HCHK(h, hipStreamSynchronize(h->stream));
    for (void* p : h->load_allocs) (void)hipFree(p);
    h->load_allocs.clear();
    I.n_load = 0;
    I.lnode = I.lrec = nullptr;
    I.lm = I.lwin = I.fgrav = nullptr;
    auto lalloc = [&](size_t bytes, void** p) -> hipError_t {
        hipError_t e = hipMalloc(p, bytes < 256 ? 256 : bytes);
        if (e == hipSuccess) h->load_allocs.push_back(*p);
        return e;
    };
    if (nf > 0) {
        void *pn, *pr, *pm, *pw;
        HCHK(h, lalloc((size_t)nf * 4, &pn));
        HCHK(h, lalloc((size_t)nf * 4, &pr));
        HCHK(h, lalloc((size_t)nf * 24, &pm));
        HCHK(h, lalloc((size_t)nf * 16, &pw));
        HCHK(h, hipMemcpy(pn, lnode.data(), (size_t)nf * 4, hipMemcpyHostToDevice));
        HCHK(h, hipMemcpy(pr, lrec.data(), (size_t)nf * 4, hipMemcpyHostToDevice));
        HCHK(h, hipMemcpy(pm, lm.data(), (size_t)nf * 24, hipMemcpyHostToDevice));
        HCHK(h, hipMemcpy(pw, lwin.data(), (size_t)nf * 16, hipMemcpyHostToDevice));
        I.lnode = (const int32_t*)pn;
        I.lrec = (const int32_t*)pr;
        I.lm = (const double*)pm;
        I.lwin = (const double*)pw;
        I.n_load = (int32_t)nf;
    }
    const int32_t gp = L ? L->gravity_pos : -1;
    I.grav_pos = (gp >= 0 && gp <= nf) ? gp : (int32_t)nf;
    HCHK(h, hipMemcpy((void*)I.ninfo, ninfo.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    if (grav != 0.0) {
        void* pg;
        HCHK(h, lalloc((size_t)N * 8, &pg));
        // the partial-force buffer (>= nen*Ne doubles) is free until the first step
        hipError_t e = mpsk::launch_gravity(I, grav, I.part, (double*)pg, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("gravity load: %s", hipGetErrorString(e)));
        I.fgrav = (const double*)pg;
    }
    return MPS_OK;
}

extern "C" int mps_external_force(mps_handle h, double t, double* fext) {
    if (!h || !fext) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    hipError_t e = mpsk::launch_fext(h->I, t, h->scratch_f, h->stream);
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("external force: %s", hipGetErrorString(e)));
    HCHK(h, hipMemcpyAsync(fext, h->scratch_f, 3 * (size_t)h->I.N * 8, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

extern "C" int mps_get_state(mps_handle h, double* u, double* v, double* a) {
    if (!h) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    const size_t bytes = 3 * (size_t)h->I.N * 8;
    if (u) HCHK(h, hipMemcpy(u, h->I.u, bytes, hipMemcpyDeviceToHost));
    if (v) HCHK(h, hipMemcpy(v, h->I.v, bytes, hipMemcpyDeviceToHost));
    if (a) HCHK(h, hipMemcpy(a, h->I.a, bytes, hipMemcpyDeviceToHost));
    return MPS_OK;
}

__global__ void k_pick(const double* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                       double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

extern "C" int mps_get_dofs(mps_handle h, int32_t which, int64_t n, const int32_t* dof, double* out) {
    if (!h || which < 0 || which > 2 || n < 0 || (n > 0 && (!dof || !out))) return MPS_ERR_ARG;
    if (n == 0) return MPS_OK;
    const int64_t ndof = 3 * (int64_t)h->I.N;
    for (int64_t i = 0; i < n; i++)
        if (dof[i] < 0 || dof[i] >= ndof) return h->fail(MPS_ERR_ARG, fmt("mps_get_dofs: dof %d outside [0, %lld)", dof[i], (long long)ndof));
    (void)hipSetDevice(h->device);
    if (n > h->pick_cap) {
        HCHK(h, hipStreamSynchronize(h->stream));
        for (void* p : h->load_allocs) (void)hipFree(p);
    h->load_allocs.clear();
    if (h->pick_idx) (void)hipFree(h->pick_idx);
        if (h->pick_out) (void)hipFree(h->pick_out);
        h->pick_idx = nullptr;
        h->pick_out = nullptr;
        h->pick_cap = 0;
        HCHK(h, hipMalloc(&h->pick_idx, (size_t)n * 4));
        HCHK(h, hipMalloc(&h->pick_out, (size_t)n * 8));
        h->pick_cap = n;
    }
    const double* src = which == 0 ? h->I.u : (which == 1 ? h->I.v : h->I.a);
    HCHK(h, hipMemcpyAsync(h->pick_idx, dof, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pick, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, src, h->pick_idx, n, h->pick_out);
    HCHK(h, hipGetLastError());
    HCHK(h, hipMemcpyAsync(out, h->pick_out, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

// Event pair around one launch when profiling; kinds: 0 element, 1 node, 2 begin, 3 energy.
struct TimedLaunch {
    mps_solver_s* h;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(mps_solver_s* hh, int k) : h(hh), kind(k) {
        if (!h->profiling) return;
        if (h->ev_used + 2 > h->evpool.size()) {
            for (int i = 0; i < 64; i++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                h->evpool.push_back(e);
            }
        }
        a = h->evpool[h->ev_used];
        b = h->evpool[h->ev_used + 1];
        h->ev_used += 2;
        (void)hipEventRecord(a, h->stream);
    }
    void done() {
        if (a) (void)hipEventRecord(b, h->stream);
    }
};

// Accumulate the recorded launch times (call after the stream is synchronised).
static void collect_times(mps_solver_s* h, const std::vector<int>& kinds) {
    for (size_t i = 0; i < kinds.size() && 2 * i + 1 < h->ev_used; i++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->evpool[2 * i], h->evpool[2 * i + 1]) == hipSuccess) {
            h->kt_ms[kinds[i]] += ms;
            h->kt_n[kinds[i]] += 1;
        }
    }
    h->ev_used = 0;
}

// Launch sequence of `count` steps starting at h->next_step; the first step is begun by
// the standalone begin kernel unless it already is; the last node pass is unfused.
static hipError_t enqueue_steps(mps_solver_s* h, long long count, std::vector<int>* kinds) {
    const MpsInst& I = h->I;
    const int every = h->cfg.energy_every;
    hipError_t e = hipSuccess;
    for (long long i = 0; i < count && e == hipSuccess; i++) {
        const long long k = h->next_step;
        if (!h->step_begun) {
            TimedLaunch t(h, 2);
            e = mpsk::launch_begin(I, h->stream);
            t.done();
            if (t.a) kinds->push_back(2);
            if (e != hipSuccess) break;
        }
        {
            TimedLaunch t(h, 0);
            e = mpsk::launch_elem(I, 0, h->stream);
            t.done();
            if (t.a) kinds->push_back(0);
        }
        if (e != hipSuccess) break;
        const bool want_e = every > 0 && (k % every) == 0;
        const bool last = (i == count - 1);
        const bool fuse = !want_e && !last;
        {
            TimedLaunch t(h, 1);
            e = mpsk::launch_node(I, fuse ? 1 : 0, h->stream);
            t.done();
            if (t.a) kinds->push_back(1);
        }
        if (e != hipSuccess) break;
        if (want_e) {
            TimedLaunch t(h, 3);
            e = mpsk::launch_energy(I, h->stream);
            t.done();
            if (t.a) kinds->push_back(3);
        }
        h->next_step = k + 1;
        h->step_begun = fuse;
    }
    return e;
}

static int read_ctl(mps_solver_s* h, MpsCtl* c) {
    HCHK(h, hipMemcpy(c, h->I.ctl, sizeof(MpsCtl), hipMemcpyDeviceToHost));
    return MPS_OK;
}

static void fill_report(mps_solver_s* h, const MpsCtl& c, mps_report* rep) {
    long long done = c.A - 1;
    int status = c.status;
    if (c.nan_step != LLONG_MAX) {
        status = MPS_ERR_NAN;
        done = c.nan_step;
    } else if (status == MPS_RUNNING && h->I.time_mode == MPS_TIME_FIXED && done >= h->I.n_steps) {
        status = MPS_DONE;
    }
    if (rep) {
        rep->steps = done;
        const MpsStepSlot& s = c.slot[done & 1];
        rep->time = done > 0 ? s.time : h->I.t0;
        rep->dt = done > 0 ? s.dt : c.slot[0].dt;
        rep->status = status;
        rep->nonfinite = c.nonfinite;
    }
}

static int status_code(int st) { return st < 0 ? st : MPS_OK; }

extern "C" int mps_run(mps_handle h, int64_t max_steps, mps_report* rep) {
    if (!h) return MPS_ERR_ARG;
    if (max_steps < 0) return h->fail(MPS_ERR_ARG, "mps_run: max_steps < 0");
    (void)hipSetDevice(h->device);
    MpsCtl c;
    int rc = read_ctl(h, &c);
    if (rc) return rc;
    mps_report r{};
    fill_report(h, c, &r);
    if (r.status != MPS_RUNNING || max_steps == 0) {
        r.device_ms = 0.0;
        if (rep) *rep = r;
        return status_code(r.status);
    }
    if (!h->primed) {
        // dt of step 1 uses uK = u0 only when t0 < 1e-12 (solverExplicit.cpp:109,176-178)
        const double* uK = (h->I.t0 < 1E-12) ? h->I.u : nullptr;
        hipError_t e = mpsk::launch_dt(h->I, uK, h->I.dtred + MPS_DT_SLOTS * MPS_DT_STRIDE, h->stream);  // parity 1, slot 0
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("first dt: %s", hipGetErrorString(e)));
        h->primed = true;
    }
    HCHK(h, hipEventRecord(h->ev0, h->stream));
    long long remaining = max_steps;
    if (h->I.time_mode == MPS_TIME_FIXED) remaining = std::min<long long>(remaining, h->I.n_steps - (h->next_step - 1));
    const long long chunk = 256;
    float total_ms = 0.f;
    while (remaining > 0) {
        long long n = std::min(remaining, chunk);
        std::vector<int> kinds;
        hipError_t e = hipSuccess;
        if (h->use_graphs && !h->profiling && h->cfg.energy_every == 0 && !h->step_begun) {
            // every chunk of n steps is the same launch sequence: replay it from a captured graph
            hipGraphExec_t ge = nullptr;
            for (auto& g : h->graphs)
                if (g.first == n) ge = g.second;
            if (!ge) {
                hipGraph_t gr = nullptr;
                const long long k0 = h->next_step;
                e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) e = enqueue_steps(h, n, &kinds);
                hipError_t e2 = hipStreamEndCapture(h->stream, &gr);
                if (e == hipSuccess) e = e2;
                if (e == hipSuccess) e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
                if (gr) (void)hipGraphDestroy(gr);
                h->next_step = k0;  // capture does not execute; replay below advances the host mirror
                h->step_begun = false;
                if (e == hipSuccess) h->graphs.emplace_back(n, ge);
            }
            if (e == hipSuccess) e = hipGraphLaunch(ge, h->stream);
            if (e == hipSuccess) {
                h->next_step += n;
                h->step_begun = false;
            }
        } else {
            e = enqueue_steps(h, n, &kinds);
        }
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("step launch: %s", hipGetErrorString(e)));
        remaining -= n;
        HCHK(h, hipEventRecord(h->ev1, h->stream));
        HCHK(h, hipEventSynchronize(h->ev1));
        if (h->profiling) collect_times(h, kinds);
        rc = read_ctl(h, &c);
        if (rc) return rc;
        fill_report(h, c, &r);
        if (r.status != MPS_RUNNING) break;
    }
    HCHK(h, hipEventElapsedTime(&total_ms, h->ev0, h->ev1));
    r.device_ms = total_ms;
    if (rep) *rep = r;
    return status_code(r.status);
}

extern "C" int mps_element_outputs(mps_handle h, double* out) {
    if (!h || !out) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    const size_t bytes = (size_t)h->I.Ne * 21 * 8;
    if (!h->eout) HCHK(h, hipMalloc(&h->eout, bytes));
    hipError_t e = mpsk::launch_elem_outputs(h->I, h->eout, h->stream);
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("element outputs: %s", hipGetErrorString(e)));
    HCHK(h, hipMemcpyAsync(out, h->eout, bytes, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

extern "C" int mps_get_history(mps_handle h, int64_t first, int64_t count, double* time, double* dt, double* e_def,
                               double* e_kin, double* volume) {
    if (!h) return MPS_ERR_ARG;
    const int cap = h->I.hist_cap;
    if (first < 0 || count < 0) return h->fail(MPS_ERR_ARG, "mps_get_history: negative range");
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> buf(5 * (size_t)cap);
    HCHK(h, hipMemcpy(buf.data(), h->I.hist, buf.size() * 8, hipMemcpyDeviceToHost));
    double* outs[5] = {time, dt, e_def, e_kin, volume};
    for (int64_t i = 0; i < count; i++) {
        size_t idx = (size_t)((first + i) % cap);
        for (int j = 0; j < 5; j++)
            if (outs[j]) outs[j][i] = buf[(size_t)j * cap + idx];
    }
    return MPS_OK;
}

extern "C" int mps_get_mass(mps_handle h, double* mass) {
    if (!h || !mass) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    HCHK(h, hipMemcpy(mass, h->I.mass, (size_t)h->I.N * 8, hipMemcpyDeviceToHost));
    return MPS_OK;
}

extern "C" int mps_internal_force(mps_handle h, const double* u, double* fint) {
    if (!h || !u || !fint) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    const size_t bytes = 3 * (size_t)h->I.N * 8;
    HCHK(h, hipMemcpyAsync(h->scratch_u, u, bytes, hipMemcpyHostToDevice, h->stream));
    MpsInst I = h->I;
    I.u = h->scratch_u;
    I.ctl = nullptr;
    hipError_t e = mpsk::launch_elem(I, 1, h->stream);
    if (e == hipSuccess) e = mpsk::launch_fint(I, h->scratch_f, h->stream);
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("internal force: %s", hipGetErrorString(e)));
    HCHK(h, hipMemcpyAsync(fint, h->scratch_f, bytes, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

extern "C" int mps_critical_dt(mps_handle h, const double* u, double* dt) {
    if (!h || !dt) return MPS_ERR_ARG;
    (void)hipSetDevice(h->device);
    const size_t bytes = 3 * (size_t)h->I.N * 8;
    if (u) HCHK(h, hipMemcpyAsync(h->scratch_u, u, bytes, hipMemcpyHostToDevice, h->stream));
    unsigned long long inf = 0x7FF0000000000000ull, bits = 0;
    HCHK(h, hipMemcpyAsync(h->scratch_dt, &inf, 8, hipMemcpyHostToDevice, h->stream));
    hipError_t e = mpsk::launch_dt(h->I, u ? h->scratch_u : nullptr, h->scratch_dt, h->stream);
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("critical dt: %s", hipGetErrorString(e)));
    HCHK(h, hipMemcpyAsync(&bits, h->scratch_dt, 8, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    double m;
    memcpy(&m, &bits, 8);
    *dt = m * h->I.scale;
    return MPS_OK;
}

extern "C" int mps_set_graphs(mps_handle h, int32_t enable) {
    if (!h) return MPS_ERR_ARG;
    h->use_graphs = enable != 0;
    return MPS_OK;
}

extern "C" int mps_profile(mps_handle h, int32_t enable) {
    if (!h) return MPS_ERR_ARG;
    h->profiling = enable != 0;
    for (int i = 0; i < 4; i++) {
        h->kt_ms[i] = 0.0;
        h->kt_n[i] = 0;
    }
    return MPS_OK;
}

extern "C" int mps_kernel_times(mps_handle h, double* ms_total, int64_t* launches) {
    if (!h) return MPS_ERR_ARG;
    for (int i = 0; i < 4; i++) {
        if (ms_total) ms_total[i] = h->kt_ms[i];
        if (launches) launches[i] = h->kt_n[i];
    }
    return MPS_OK;
}

extern "C" int mps_destroy(mps_handle h) {
    if (!h) return MPS_ERR_ARG;
    release(h);
    delete h;
    return MPS_OK;
}

extern "C" const char* mps_last_error(mps_handle h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

extern "C" int mps_device_info(int32_t device, char* buf, int32_t len) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return MPS_ERR_HIP;
    if (buf && len > 0)
        snprintf(buf, len, "%s gcn=%s CUs=%d HBM=%.1fGiB clock=%dMHz", p.name, p.gcnArchName, p.multiProcessorCount,
                 p.totalGlobalMem / 1073741824.0, p.clockRate / 1000);
    return MPS_OK;
}
