// mps_api.hip -- C-ABI of libmps_gpu: single-solve path (see include/mps_gpu.h).
//
// Host responsibilities: validate inputs, build the node->(element,slot) CSR, the
// Dirichlet records and per-element material constants, upload, run the set-up
// kernels, then drive the step kernels (optionally through a captured hipGraph)
// without any per-step host round trip.  Every failure returns a status code;
// nothing here calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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
#include "mps_device.h"
#include "mps_kernels.h"

// Staging of mps_get_dofs_static: MPS_DOFS_STATIC_MAX DOF indices (int32), then as many gathered values
// (double).  Every device the library runs on holds one copy until the process ends.  Host code never uses
// this name as an address: each device's copy is reached through the pointer resolved once per device by
// hipGetSymbolAddress in mps_create (scratch_base).
#define MPS_STATIC_IDX_BYTES (((size_t)MPS_DOFS_STATIC_MAX * 4 + 255) / 256 * 256)
static __device__ __align__(256) unsigned char scratch[MPS_STATIC_IDX_BYTES + (size_t)MPS_DOFS_STATIC_MAX * 8];

namespace {

std::mutex g_err_mu;
std::string g_err;  // last mps_create failure

// Device address of `scratch` on each device, resolved once with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;
std::vector<hipError_t> g_scratch_err;

void resolve_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) {
        g_scratch_ptr.resize(device + 1, nullptr);
        g_scratch_err.resize(device + 1, hipErrorNotInitialized);
    }
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = p;
}

void* scratch_base(int device, hipError_t* why) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= (int)g_scratch_ptr.size()) {
        *why = hipErrorNotInitialized;
        return nullptr;
    }
    *why = g_scratch_err[device];
    return g_scratch_ptr[device];
}

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

// Element storage order of hex8 meshes: ord[s] = input id of the element stored at position s (empty:
// input order).  The elements are sorted along a space-filling curve through their centroids, so that the
// MPS_PATCH consecutive elements of a patch form a compact block with few distinct nodes.  Each distinct
// node of a patch costs one partial force, which the element kernel writes and the node kernel reads: a
// perturbed 100^3 hex box numbered x-fastest has 2.83 partials per element in input order and 1.66 in
// this order.
// Centroids are measured along each axis in cells of the mean element size there: the mean extent of the
// elements' bounding boxes, scaled so that the three sizes multiply to the mean element volume (box
// volume / Ne; the bounding boxes of perturbed elements overstate it).  Key bits x0 x1 x2 x3 y0 z0 y1 z1,
// then x, y, z interleaved: a wavefront's 64 elements are a 16x2x2 block of cells and a patch 16x4x4, so
// each corner's node gathers stay runs of about 16 nodes.  Ties keep the input order.
// Tet4 meshes keep the input order: the tet element kernel, which stages each patch's distinct nodes and
// sums each node's contributions in one work-item, got slower on compact patches (longer sums per node)
// than the node kernel gained.  MPS_ELEM_ORDER=input in the environment keeps the input order for hex8 too.
static void element_order(const double* X, int64_t N, const int32_t* conn, int64_t Ne, int nen,
                          std::vector<int32_t>* ord) {
    ord->clear();
    const char* env = getenv("MPS_ELEM_ORDER");
    if ((env && strcmp(env, "input") == 0) || Ne < 2 || nen != 8) return;
    double lo[3], hi[3], h[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; d++) {
        lo[d] = hi[d] = X[d * N];
        for (int64_t n = 1; n < N; n++) {
            lo[d] = std::min(lo[d], X[d * N + n]);
            hi[d] = std::max(hi[d], X[d * N + n]);
        }
    }
    // one pass over the elements: centroid (kept) and bounding-box extent (summed)
    std::vector<double> cen(3 * (size_t)Ne);
    for (int64_t e = 0; e < Ne; e++) {
        const int32_t* c = conn + e * nen;
        for (int d = 0; d < 3; d++) {
            const double* x = X + d * N;
            double a = x[c[0]], b = a, sum = a;
            for (int k = 1; k < nen; k++) {
                const double v = x[c[k]];
                a = std::min(a, v);
                b = std::max(b, v);
                sum += v;
            }
            h[d] += b - a;
            cen[3 * (size_t)e + d] = sum / nen;
        }
    }
    const double emax = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const double hmax = std::max(h[0], std::max(h[1], h[2]));
    if (!(emax > 0.0) || !std::isfinite(emax) || !(hmax > 0.0) || !std::isfinite(hmax)) return;
    double q[3], vol = 1.0, hv = 1.0;
    for (int d = 0; d < 3; d++) {
        vol *= std::max(hi[d] - lo[d], 1e-6 * emax);
        h[d] = std::max(h[d], 1e-6 * hmax) / (double)Ne;
        hv *= h[d];
    }
    for (int d = 0; d < 3; d++) q[d] = h[d] * std::cbrt(vol / (double)Ne / hv);
    // key bit positions of axis d's bits 0..19: x0 x1 x2 x3 y0 z0 y1 z1 at 0..7, then x, y, z interleaved
    int kpos[3][20];
    {
        int len = 0;
        const int head[8][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 0}, {2, 0}, {1, 1}, {2, 1}};
        for (int k = 0; k < 8; k++) kpos[head[k][0]][head[k][1]] = len++;
        for (int b = 0; len < 60; b++) {
            const int want[3] = {4 + b, 2 + b, 2 + b};
            for (int d = 0; d < 3; d++)
                if (want[d] < 20) kpos[d][want[d]] = len++;
        }
    }
    // the key of a cell index: three table lookups per axis (bits 0-7, 8-15, 16-19)
    std::vector<uint64_t> tab(3 * 3 * 256, 0);
    for (int d = 0; d < 3; d++)
        for (int part = 0; part < 3; part++)
            for (int v = 0; v < 256; v++)
                for (int bit = 0; bit < 8 && part * 8 + bit < 20; bit++)
                    if ((v >> bit) & 1) tab[(d * 3 + part) * 256 + v] |= 1ull << kpos[d][part * 8 + bit];
    std::vector<uint64_t> key((size_t)Ne);
    uint64_t kmax = 0;
    for (int64_t e = 0; e < Ne; e++) {
        uint64_t k = 0;
        for (int d = 0; d < 3; d++) {
            const double t = (cen[3 * (size_t)e + d] - lo[d]) / q[d];
            const uint32_t c = t > 0.0 ? (uint32_t)std::min(t, (double)((1 << 20) - 1)) : 0u;
            k |= tab[(d * 3 + 0) * 256 + (c & 255)] | tab[(d * 3 + 1) * 256 + ((c >> 8) & 255)] |
                 tab[(d * 3 + 2) * 256 + (c >> 16)];
        }
        key[(size_t)e] = k;
        kmax = std::max(kmax, k);
    }
    // sort by (key, input id): both packed into one word when they fit (an LSD radix sort on 8-bit digits),
    // else a comparison sort of the pairs
    int ibits = 1, kbits = 0;
    while (((int64_t)1 << ibits) < Ne) ibits++;
    while (kbits < 64 && (kmax >> kbits) != 0) kbits++;
    ord->resize((size_t)Ne);
    if (kbits + ibits <= 64) {
        std::vector<uint64_t> w((size_t)Ne), tmp((size_t)Ne);
        for (int64_t e = 0; e < Ne; e++) w[(size_t)e] = (key[(size_t)e] << ibits) | (uint64_t)e;
        for (int shift = 0; shift < kbits + ibits; shift += 8) {
            size_t cnt[257] = {0};
            for (uint64_t v : w) cnt[((v >> shift) & 255) + 1]++;
            for (int i = 0; i < 256; i++) cnt[i + 1] += cnt[i];
            for (uint64_t v : w) tmp[cnt[(v >> shift) & 255]++] = v;
            w.swap(tmp);
        }
        const uint64_t imask = ((uint64_t)1 << ibits) - 1;
        for (int64_t s = 0; s < Ne; s++) (*ord)[(size_t)s] = (int32_t)(w[(size_t)s] & imask);
    } else {
        std::vector<std::pair<uint64_t, int32_t>> kp((size_t)Ne);
        for (int64_t e = 0; e < Ne; e++) kp[(size_t)e] = {key[(size_t)e], (int32_t)e};
        std::sort(kp.begin(), kp.end());
        for (int64_t s = 0; s < Ne; s++) (*ord)[(size_t)s] = kp[(size_t)s].second;
    }
    bool same = true;
    for (int64_t s = 0; s < Ne && same; s++) same = (*ord)[(size_t)s] == (int32_t)s;
    if (same) ord->clear();
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
    // CSR node -> element slots, in ascending input element, then slot: the lumped mass and the gravity
    // load add their element terms in the reference's order.  The entries address the stored slots.
    hp->rowptr.assign(N + 1, 0);
    for (int64_t k = 0; k < Ne * nen; k++) {
        int32_t n = mesh->conn[k];
        if (n < 0 || n >= N) {
            *err = fmt("element %lld references node %d outside [0,%lld)", (long long)(k / nen), n, (long long)N);
            return MPS_ERR_ARG;
        }
        hp->rowptr[n + 1]++;
    }
    for (int64_t n = 0; n < N; n++) hp->rowptr[n + 1] += hp->rowptr[n];
    element_order(mesh->X, N, mesh->conn, Ne, (int)nen, &hp->eorig);
    const std::vector<int32_t>& ord = hp->eorig;
    std::vector<int32_t> pos;  // stored position of each input element
    hp->conn.resize(Ne * nen);
    if (ord.empty()) {
        hp->conn.assign(mesh->conn, mesh->conn + Ne * nen);
    } else {
        pos.resize((size_t)Ne);
        for (int64_t s = 0; s < Ne; s++) {
            pos[(size_t)ord[(size_t)s]] = (int32_t)s;
            memcpy(&hp->conn[s * nen], mesh->conn + (int64_t)ord[(size_t)s] * nen, nen * 4);
        }
    }
    hp->col.resize(Ne * nen);
    {
        std::vector<int32_t> fill(hp->rowptr.begin(), hp->rowptr.end() - 1);
        for (int64_t k = 0; k < Ne * nen; k++) {
            const int64_t e = k / nen, slot = pos.empty() ? k : (int64_t)pos[(size_t)e] * nen + (k - e * nen);
            hp->col[fill[mesh->conn[k]]++] = (int32_t)slot;
        }
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
        // per patch: its distinct nodes (marked with the patch index), sorted; each node's local index; the
        // contributions counted per local node, then placed in ascending slot order within each node's group
        std::vector<int32_t> mark((size_t)N, -1), lidx((size_t)N), dn, start((size_t)(P * nen + 1));
        dn.reserve((size_t)(P * nen));
        for (int64_t p = 0; p < np; p++) {
            const int64_t e0 = p * P, m = (std::min(e0 + P, Ne) - e0) * nen;
            const int32_t* c = hp->conn.data() + e0 * nen;
            dn.clear();
            for (int64_t q = 0; q < m; q++)
                if (mark[(size_t)c[q]] != (int32_t)p) {
                    mark[(size_t)c[q]] = (int32_t)p;
                    dn.push_back(c[q]);
                }
            std::sort(dn.begin(), dn.end());
            const int64_t U = (int64_t)dn.size();
            start[0] = 0;
            for (int64_t j = 0; j < U; j++) {
                lidx[(size_t)dn[(size_t)j]] = (int32_t)j;
                start[(size_t)j + 1] = 0;
            }
            for (int64_t q = 0; q < m; q++) start[(size_t)lidx[(size_t)c[q]] + 1]++;
            for (int64_t j = 0; j < U; j++) {
                start[(size_t)j + 1] += start[(size_t)j];
                hp->prow.push_back((uint16_t)start[(size_t)j]);
            }
            hp->prow.push_back((uint16_t)m);
            for (int64_t q = 0; q < m; q++) hp->pcol[(size_t)(e0 * nen + start[(size_t)lidx[(size_t)c[q]]]++)] = (uint16_t)q;
            pnode.insert(pnode.end(), dn.begin(), dn.end());
            hp->poff[p + 1] = hp->poff[p] + (int32_t)U;
        }
        hp->n_part = hp->poff[np];
        hp->nrow.assign(N + 1, 0);
        for (int32_t n : pnode) hp->nrow[n + 1]++;
        for (int64_t n = 0; n < N; n++) hp->nrow[n + 1] += hp->nrow[n];
        hp->ncol.resize(pnode.size());
        std::vector<int32_t> fill(hp->nrow.begin(), hp->nrow.end() - 1);
        for (size_t g = 0; g < pnode.size(); g++) hp->ncol[fill[pnode[g]]++] = (int32_t)g;
        hp->pnd = std::move(pnode);
    }
    // materials
    hp->per_elem = mat->per_element != 0;
    if (hp->per_elem) {
        hp->mat.resize(2 * Ne);
        hp->cel.resize(Ne);
        hp->rho.resize(Ne);
        for (int64_t e = 0; e < Ne; e++) {  // stored order
            const int64_t ei = ord.empty() ? e : ord[(size_t)e];
            double E = mat->E[ei], nu = mat->nu[ei], rho = mat->rho[ei];
            if (!(E > 0) || !(nu > -1.0 && nu < 0.5) || !(rho > 0)) {
                *err = fmt("element %lld: invalid material E=%g nu=%g rho=%g", (long long)ei, E, nu, rho);
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
    std::vector<int32_t> epos;    // stored position of each input element (empty: input order)
    int32_t* pick_idx = nullptr;  // mps_get_dofs buffers (grow-only, freed with the handle)
    double* pick_out = nullptr;
    int64_t pick_cap = 0;
    std::vector<void*> load_allocs;  // mps_set_loads buffers (replaced by the next call)
    double* uprev_buf = nullptr;     // previous-step displacement for bulk viscosity (in allocs)
    double* fint_part = nullptr;     // partial forces of mps_internal_force (allocated on first use, in allocs)
    bool part_valid = false;         // I.part holds the forces of the current u (false: set-up scratch)
    bool primed = false;       // prologue (first dt) done
    long long next_step = 1;   // host mirror of ctl->A
    bool step_begun = false;   // step next_step already predicted on the device
    struct Graph {
        long long n, phase;  // steps, first step modulo energy_every (the energy launches of the chunk)
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;  // captured chunks
    bool use_graphs = true;
    // optional kernel timing (HIP events around each element / node launch)
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    size_t ev_used = 0;
    double kt_ms[4] = {0, 0, 0, 0};   // element ms, node ms, begin ms, energy ms
    long long kt_n[4] = {0, 0, 0, 0};
    // output snapshots (mps_snapshot_*): two slots, each a device staging copy and a pinned host copy of
    // u, v, a and the element outputs, filled on the solver stream and copied out on copy_stream
    struct Snap {
        double* dev = nullptr;
        double* host = nullptr;
        hipEvent_t staged = nullptr, copied = nullptr;
        std::atomic<int> busy{0};  // begun and not yet released (mps_snapshot_release may run on another thread)
        int with_elem = 0;
    } snap[2];
    hipStream_t copy_stream = nullptr;
    std::vector<void*> series_allocs;  // mps_set_dof_series buffers (replaced by the next call)
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
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    for (void* p : h->load_allocs) (void)hipFree(p);
    h->load_allocs.clear();
    for (void* p : h->series_allocs) (void)hipFree(p);
    h->series_allocs.clear();
    if (h->pick_idx) (void)hipFree(h->pick_idx);
    if (h->pick_out) (void)hipFree(h->pick_out);
    if (h->eout) (void)hipFree(h->eout);
    h->pick_idx = nullptr;
    h->pick_out = nullptr;
    h->eout = nullptr;
    if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
    for (auto& S : h->snap) {
        if (S.dev) (void)hipFree(S.dev);
        if (S.host) (void)hipHostFree(S.host);
        if (S.staged) (void)hipEventDestroy(S.staged);
        if (S.copied) (void)hipEventDestroy(S.copied);
        S.dev = S.host = nullptr;
        S.staged = S.copied = nullptr;
    }
    if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
    h->copy_stream = nullptr;
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

static int upload_instance(mps_solver_s* h, const HostPrep& hp) {
    MpsInst& I = h->I;
    const size_t N = hp.N, Ne = hp.Ne, nen = hp.nen;
    I.N = hp.N;
    // node gathers of the hex8 element kernels by 32-bit byte offsets from scalar bases when every offset fits;
    // the environment variable MPS_GATHER32=0 keeps 64-bit addresses (same loads, same results bit for bit)
    {
        const char* g32 = getenv("MPS_GATHER32");
        I.off32 = (hp.N <= MPS_OFF32_NODES && !(g32 && g32[0] == '0')) ? 1 : 0;
    }
    // stable-dt skip of the FP64 hex8 element kernels: an element whose dt is certainly more than 1/64 above the
    // previous step's min skips the strict length (the certificate keeps the min exact).  The environment
    // variable MPS_DT_SKIP=0 turns it off; another number replaces the margin 1/64 (a negative one makes the
    // certificate fail, which exercises the fallback).  Same results, bit for bit, either way.
    {
        const char* ds = getenv("MPS_DT_SKIP");
        I.dt_skip_on = (ds && strcmp(ds, "0") == 0) ? 0 : 1;
        I.dt_skip_margin = (ds && I.dt_skip_on) ? atof(ds) : 1.0 / 64;
    }
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
    double *X, *u, *v, *a, *mass, *l0, *matp = nullptr, *cel = nullptr, *rho = nullptr, *bcval, *hist;
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
    // prow and pcol carry 8 entries of padding: the element kernels read them in 8-entry pieces
    HCHK(h, dalloc(h, &prow, (NP + NPA + 8) * 2));
    HCHK(h, dalloc(h, &pcol, (Ne * nen + 8) * 2));
    HCHK(h, dalloc(h, &nrow, (N + 1) * 4));
    HCHK(h, dalloc(h, &ncol, NP * 4));
    HCHK(h, dalloc(h, &ctl, sizeof(MpsCtl)));
    unsigned long long* dtred;
    HCHK(h, dalloc(h, &dtred, (2 * MPS_DT_SLOTS * MPS_DT_STRIDE + MPS_DT_COUNT_WORDS) * 8));
    HCHK(h, hipMemsetAsync(dtred + 2 * MPS_DT_SLOTS * MPS_DT_STRIDE, 0, MPS_DT_COUNT_WORDS * 8, h->stream));
    HCHK(h, dalloc(h, &hist, MPS_HIST_COLS * (size_t)cap * 8));
    if (hp.cfg.energy_every > 0) {  // energy block sums (element phase, node kernel by parity), support forces
        const size_t node_blocks = (N + 255) / 256;
        HCHK(h, dalloc(h, &I.epart, (3 * NPA + 2 * MPS_NODE_PARTS * node_blocks) * 8));
        HCHK(h, dalloc(h, &I.rprev, hp.bcval.size() * 8));
        HCHK(h, hipMemsetAsync(I.rprev, 0, hp.bcval.size() * 8, h->stream));
    }
    I.energy_every = hp.cfg.energy_every > 0 ? hp.cfg.energy_every : 0;
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
    if (!hp.eorig.empty()) {
        int32_t* eorig;
        HCHK(h, dalloc(h, &eorig, Ne * 4));
        HCHK(h, hipMemcpyAsync(eorig, hp.eorig.data(), Ne * 4, hipMemcpyHostToDevice, h->stream));
        I.eorig = eorig;
        h->epos.resize(Ne);
        for (size_t s = 0; s < Ne; s++) h->epos[(size_t)hp.eorig[s]] = (int32_t)s;
    }
    HCHK(h, hipMemcpyAsync(rowptr, hp.rowptr.data(), (N + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(col, hp.col.data(), Ne * nen * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(poff, hp.poff.data(), (NPA + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(prow, hp.prow.data(), (NP + NPA) * 2, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(pcol, hp.pcol.data(), Ne * nen * 2, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemsetAsync(prow + NP + NPA, 0, 16, h->stream));
    HCHK(h, hipMemsetAsync(pcol + Ne * nen, 0, 16, h->stream));
    HCHK(h, hipMemcpyAsync(nrow, hp.nrow.data(), (N + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(ncol, hp.ncol.data(), NP * 4, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(ninfo, hp.ninfo.data(), N * 4, hipMemcpyHostToDevice, h->stream));
    // tet4 node staging tables: local index of every element slot, and the largest patch node count,
    // when the staging buffer and the force slots fit the 64 KB of LDS a workgroup may use.  The
    // environment variable MPS_TET_STAGING=0 keeps the per-element gathers (same results, bit for bit).
    const char* stage_env = getenv("MPS_TET_STAGING");
    if (nen == 4 && !(stage_env && stage_env[0] == '0')) {
        std::vector<uint16_t> plx((size_t)(Ne * nen));
        int32_t pu = 0;
        for (int64_t p = 0; p < (int64_t)NPA; p++) {
            const int32_t b = hp.poff[p], U = hp.poff[p + 1] - b;
            const uint16_t* pr = hp.prow.data() + b + p;
            const uint16_t* pc = hp.pcol.data() + p * MPS_PATCH * nen;
            for (int32_t j = 0; j < U; j++)
                for (int q = pr[j]; q < pr[j + 1]; q++) plx[(size_t)p * MPS_PATCH * nen + pc[q]] = (uint16_t)j;
            pu = std::max(pu, U);
        }
        // node staging and FP64 force slots (one shared space), contribution list, and the kernel's static
        // block-min slots (with room to spare)
        const size_t lds = std::max<size_t>(48 * (size_t)pu, 12 * MPS_PATCH * 8) + 8 * MPS_PATCH + 64;
        if (lds <= 65536) {
            int32_t* pnd;
            uint16_t* plxd;
            HCHK(h, dalloc(h, &pnd, NP * 4));
            HCHK(h, dalloc(h, &plxd, Ne * nen * 2));
            HCHK(h, hipMemcpyAsync(pnd, hp.pnd.data(), NP * 4, hipMemcpyHostToDevice, h->stream));
            HCHK(h, hipMemcpyAsync(plxd, plx.data(), Ne * nen * 2, hipMemcpyHostToDevice, h->stream));
            // the host vector must outlive the asynchronous copy
            HCHK(h, hipStreamSynchronize(h->stream));
            I.pnd = pnd;
            I.plx = plxd;
            I.pu_max = pu;
        }
    }
    HCHK(h, hipMemcpyAsync(bcval, hp.bcval.data(), hp.bcval.size() * 8, hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemsetAsync(hist, 0xFF, MPS_HIST_COLS * (size_t)cap * 8, h->stream));  // NaN
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
    c.dt_skip = HUGE_VAL;  // the first step computes every element's dt
    c.skip_hold = 0;
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
    resolve_scratch(h->device);  // once per device, with the device current; a failure surfaces in its user
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
    h->part_valid = false;
    // the handed-over u is also the previous state of the first step (solverExplicit.cpp:149)
    if (h->I.uprev) HCHK(h, hipMemcpyAsync(h->I.uprev, h->I.u, bytes, hipMemcpyDeviceToDevice, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

extern "C" int mps_set_viscosity(mps_handle h, int32_t type, double c1, double cl, double elem_size, double c0) {
    if (!h) return MPS_ERR_ARG;
    if (h->primed || h->next_step != 1)
        return h->fail(MPS_ERR_STATE, "mps_set_viscosity: only valid before the first step");
    if (type < MPS_VISC_NONE || type > MPS_VISC_DEV) return h->fail(MPS_ERR_ARG, "mps_set_viscosity: unknown type");
    if (!std::isfinite(c1) || !std::isfinite(cl) || !std::isfinite(elem_size) || !std::isfinite(c0))
        return h->fail(MPS_ERR_ARG, "mps_set_viscosity: parameters must be finite");
    MpsInst& I = h->I;
    if (type != MPS_VISC_BULK) {  // the deviatoric variant adds no force
        I.uprev = nullptr;
        return MPS_OK;
    }
    (void)hipSetDevice(h->device);
    const size_t bytes = 3 * (size_t)I.N * 8;
    if (!h->uprev_buf) HCHK(h, dalloc(h, &h->uprev_buf, bytes));
    HCHK(h, hipMemcpyAsync(h->uprev_buf, I.u, bytes, hipMemcpyDeviceToDevice, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    I.uprev = h->uprev_buf;
    I.visc_c1 = c1;
    I.visc_cl = cl;
    I.visc_h = elem_size;
    I.visc_c0 = c0;
    return MPS_OK;
}

// ---------------------------------------------------------------------------
// Neumann loads: host set-up
// ---------------------------------------------------------------------------
namespace {

// Faces of getSubElement: hex (classLinearHexahedra.cpp:30-85), tet (classLinearTetrahedra.cpp:30-55).
const int kHexFace[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
const int kTetFace[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {0, 2, 3}};

// multGeneralMatrices: C[m x ll] = A[m x kk] B[kk x ll], k innermost from zero (maths.cpp:1679-1693).
void mat_general(const double* A, int m, int kk, const double* B, int ll, double* C) {
    for (int i = 0; i < m * ll; i++) C[i] = 0.0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < ll; j++)
            for (int k = 0; k < kk; k++) C[i * ll + j] += A[i * kk + k] * B[k * ll + j];
}

// inverse, 3x3 branch (maths.cpp:2749-2767).
void inverse3(const double* A, double* R) {
    const double idet = 1.0 / mpsd::det3_strict(A);
    R[0] = idet * (A[4] * A[8] - A[5] * A[7]);
    R[3] = -idet * (A[3] * A[8] - A[5] * A[6]);
    R[6] = idet * (A[3] * A[7] - A[4] * A[6]);
    R[1] = -idet * (A[1] * A[8] - A[2] * A[7]);
    R[4] = idet * (A[0] * A[8] - A[2] * A[6]);
    R[7] = -idet * (A[0] * A[7] - A[1] * A[6]);
    R[2] = idet * (A[1] * A[5] - A[2] * A[4]);
    R[5] = -idet * (A[0] * A[5] - A[2] * A[3]);
    R[8] = idet * (A[0] * A[4] - A[1] * A[3]);
}

double norm2_3(const double* a) {  // maths.cpp:1550-1557
    double t = 0;
    for (int i = 0; i < 3; i++) t += a[i] * a[i];
    return std::sqrt(t);
}

void cross3(const double* A, const double* B, double* C) {  // geometry.cpp:180-190
    C[0] = A[1] * B[2] - A[2] * B[1];
    C[1] = A[2] * B[0] - A[0] * B[2];
    C[2] = A[0] * B[1] - A[1] * B[0];
}

// Surface Gauss points of face S (1-based) of bulk element e, as the reference builds a pressure face:
//   getn0 of the face (classLinearSquare.cpp:19-45, classLinearTriangles.cpp:19-45);
//   the face's defineGPs in the plane basis of createOrthogonalBasisOnPlane (geometry.cpp:363-389):
//     quad 2x2 points +-0.577350269, weight 1 (classLinearSquare.cpp:196-301, GPsDistribution.cpp:297-318);
//     triangle one point (0.33333333333333, 0.33333333333333), weight 0.5*1 (classLinearTriangles.cpp:194-299);
//   defineShapeOnAPoint of the bulk element at the point (classLinearHexahedra.cpp:95-172: natural coordinates
//   2 A^-1 (x - x_0) - 1, A the corner edges 0-1, 0-3, 0-4; classLinearTetrahedra.cpp:66-117: the shape at
//   the natural origin) for F and, with PRESSURE RAMP, the scatter.
// Per point: out[MPS_SGP_STRIDE] = bulk dphi (8x3), n0, w|J_s|, scatter phi (8); nodes[8] scatter nodes (-1
// where unused).  Returns the number of points, or -1 for a face outside the element.
int surface_points(const std::vector<double>& X, int64_t N, const int32_t* cn, int nen, int S, bool ramp, double* out,
                   int32_t* nodes) {
    const int nf = (nen == 8) ? 4 : 3;
    if (S < 1 || S > ((nen == 8) ? 6 : 4)) return -1;
    const int* fl = (nen == 8) ? kHexFace[S - 1] : kTetFace[S - 1];
    double xb[8][3], xf[4][3];
    for (int a = 0; a < nen; a++)
        for (int d = 0; d < 3; d++) xb[a][d] = X[(size_t)d * N + cn[a]];
    for (int k = 0; k < nf; k++)
        for (int d = 0; d < 3; d++) xf[k][d] = xb[fl[k]][d];
    double nn[3];
    for (int i = 0; i < 3; i++) {
        const int c1 = (i == 0) ? 1 : (i == 1) ? 2 : 0, c2 = (i == 0) ? 2 : (i == 1) ? 0 : 1;
        nn[i] = (xf[1][c1] - xf[0][c1]) * (xf[2][c2] - xf[0][c2]) - (xf[2][c1] - xf[0][c1]) * (xf[1][c2] - xf[0][c2]);
    }
    const double no = norm2_3(nn);
    double a[3], b[3], normal[3], xhat[3], map[6];
    for (int i = 0; i < 3; i++) {
        a[i] = xf[1][i] - xf[0][i];
        b[i] = xf[2][i] - xf[0][i];
    }
    cross3(b, a, normal);
    cross3(a, normal, xhat);
    const double magA = norm2_3(a), magX = norm2_3(xhat);
    for (int i = 0; i < 3; i++) {
        map[3 + i] = a[i] / magA;
        map[i] = xhat[i] / magX;
    }
    double tmp[12], loc[8];
    for (int k = 0; k < nf; k++)
        for (int d = 0; d < 3; d++) tmp[d * nf + k] = xf[k][d] - xf[0][d];
    mat_general(map, 2, 3, tmp, nf, loc);
    const double G = 0.577350269;
    const double sq[4][2] = {{G, G}, {-G, G}, {G, -G}, {-G, -G}};
    const int ng = (nen == 8) ? 4 : 1;
    double xyzCoor[24];
    for (int k = 0; k < nen; k++)
        for (int d = 0; d < 3; d++) xyzCoor[d * nen + k] = xb[k][d];
    for (int g = 0; g < ng; g++) {
        double ph[4] = {0, 0, 0, 0}, nd[8], Jac[4], weight;
        if (nen == 8) {
            const double x = sq[g][0], y = sq[g][1];
            ph[0] = 0.25 * (1 - x) * (1 - y);
            ph[1] = 0.25 * (1 + x) * (1 - y);
            ph[2] = 0.25 * (1 + x) * (1 + y);
            ph[3] = 0.25 * (1 - x) * (1 + y);
            nd[0] = -0.25 * (1 - y);
            nd[2] = 0.25 * (1 - y);
            nd[4] = 0.25 * (1 + y);
            nd[6] = -0.25 * (1 + y);
            nd[1] = -0.25 * (1 - x);
            nd[3] = -0.25 * (1 + x);
            nd[5] = 0.25 * (1 + x);
            nd[7] = 0.25 * (1 - x);
            weight = 1 * 1.0;
        } else {
            const double x = 0.33333333333333, y = 0.33333333333333;
            ph[0] = -1.0 * x - 1.0 * y + 1.0;
            ph[1] = 1.0 * x;
            ph[2] = 1.0 * y;
            nd[0] = -1.0 - 0.0 + 0.0;
            nd[2] = 1.0;
            nd[4] = 0.0;
            nd[1] = -0.0 - 1.0 + 0.0;
            nd[3] = 0.0;
            nd[5] = 1.0;
            weight = 0.5 * 1.0;
        }
        mat_general(loc, 2, nf, nd, 2, Jac);
        const double Js = Jac[0] * Jac[3] - Jac[2] * Jac[1];
        double x[3];
        for (int i = 0; i < 3; i++) {
            x[i] = 0;
            for (int k = 0; k < nf; k++) x[i] += xf[k][i] * ph[k];
        }
        // bulk shape at the point
        double phb[8] = {0, 0, 0, 0, 0, 0, 0, 0}, ndb[24];
        if (nen == 8) {
            double A[9], Ai[9], Xr[3], xw[3];
            const int c[3] = {1, 3, 4};
            for (int d = 0; d < 3; d++)
                for (int j = 0; j < 3; j++) A[d * 3 + j] = xb[c[j]][d] - xb[0][d];
            for (int d = 0; d < 3; d++) Xr[d] = x[d] - xb[0][d];
            inverse3(A, Ai);
            for (int i = 0; i < 3; i++) {
                double r = 0;
                for (int j = 0; j < 3; j++) r += Ai[i * 3 + j] * Xr[j];
                xw[i] = 2 * r - 1;
            }
            mpsd::hex_shape_ref(xw[0], xw[1], xw[2], phb, ndb);
        } else {  // classLinearTetrahedra::getInterpolationParameters at (0, 0, 0)
            const double q = 0, r = 0, w = 0;
            phb[0] = -1.0 * q - 1.0 * r - 1.0 * w + 1.0;
            phb[1] = 1.0 * q;
            phb[2] = 1.0 * r;
            phb[3] = 1.0 * w;
            for (int k = 0; k < 12; k++) ndb[k] = 0.0;
            ndb[0] = -1.0 - 0.0 - 0.0 + 0.0;
            ndb[3] = 1.0;
            ndb[1] = -0.0 - 1.0 - 0.0 + 0.0;
            ndb[7] = 1.0;
            ndb[2] = -0.0 - 0.0 - 1.0 + 0.0;
            ndb[11] = 1.0;
        }
        double Jb[9], iJb[9];
        mat_general(xyzCoor, 3, nen, ndb, 3, Jb);
        inverse3(Jb, iJb);
        double* o = out + (size_t)g * MPS_SGP_STRIDE;
        mat_general(ndb, nen, 3, iJb, 3, o);  // dphi = natDev * J^-1
        for (int k = 3 * nen; k < 24; k++) o[k] = 0.0;
        for (int i = 0; i < 3; i++) o[24 + i] = nn[i] / no;
        o[27] = weight * std::fabs(Js);
        for (int k = 0; k < 8; k++) {
            o[28 + k] = ramp ? phb[k] : (k < nf ? ph[k] : 0.0);
            nodes[g * 8 + k] = ramp ? (k < nen ? cn[k] : -1) : (k < nf ? cn[fl[k]] : -1);
        }
    }
    return ng;
}

struct LoadItem {
    int32_t node, block;
    int64_t seq;
    uint32_t code;
};

}  // namespace

// Blocks in input order become per-node load items: each FORCE entry, each (pressure Gauss point, scatter
// node), and gravity for every node that has other items.  Items are sorted by node, each node's in input
// order, and flagged in ninfo; gravity is summed per node once on the device.
extern "C" int mps_set_loads(mps_handle h, int32_t n_loads, const mps_load* loads) {
    if (!h) return MPS_ERR_ARG;
    if (h->primed || h->next_step != 1) return h->fail(MPS_ERR_STATE, "mps_set_loads: only valid before the first step");
    if (n_loads < 0 || (n_loads > 0 && !loads)) return h->fail(MPS_ERR_ARG, "mps_set_loads: bad block list");
    MpsInst& I = h->I;
    const int64_t N = I.N, Ne = I.Ne;
    const int nen = I.nen;
    int n_grav = 0, grav_block = -1;
    double grav = 0.0;
    int64_t n_force = 0, n_face = 0;
    for (int b = 0; b < n_loads; b++) {
        const mps_load& L = loads[b];
        const std::string at = fmt("mps_set_loads: block %d: ", b);
        if (L.n < 0) return h->fail(MPS_ERR_ARG, at + "n < 0");
        if (L.type == MPS_LOAD_FORCE) {
            if (L.n > 0 && (!L.index || !L.force)) return h->fail(MPS_ERR_ARG, at + "FORCE needs index and force");
            if (!std::isfinite(L.t0) || !std::isfinite(L.tf) || !(L.tf > L.t0))
                return h->fail(MPS_ERR_ARG, at + "FORCE window needs finite tf > t0");
            for (int k = 0; k < L.n; k++) {
                if (L.index[k] < 0 || L.index[k] >= N)
                    return h->fail(MPS_ERR_ARG, at + fmt("node %d outside [0,%lld)", L.index[k], (long long)N));
                for (int d = 0; d < 3; d++)
                    if (!std::isfinite(L.force[3 * (size_t)k + d])) return h->fail(MPS_ERR_ARG, at + "force is not finite");
            }
            n_force += L.n;
        } else if (L.type == MPS_LOAD_GRAVITY) {
            if (!std::isfinite(L.value)) return h->fail(MPS_ERR_ARG, at + "gravity is not finite");
            if (++n_grav > 1) return h->fail(MPS_ERR_ARG, at + "a second GRAVITY block: one gravity load per solve");
            grav = L.value;
            grav_block = b;
        } else if (L.type >= MPS_LOAD_PRESSURE_RAMP && L.type <= MPS_LOAD_TRACTION_INST) {
            const bool ramp = L.type == MPS_LOAD_PRESSURE_RAMP || L.type == MPS_LOAD_TRACTION_RAMP;
            const bool trac = L.type >= MPS_LOAD_TRACTION_RAMP;
            if (L.n > 0 && (!L.index || !L.face)) return h->fail(MPS_ERR_ARG, at + "surface load needs index and face");
            if (trac && (!L.force || !std::isfinite(L.force[0]) || !std::isfinite(L.force[1]) || !std::isfinite(L.force[2])))
                return h->fail(MPS_ERR_ARG, at + "TRACTION needs a finite traction vector in force[0..2]");
            if ((!trac && !std::isfinite(L.value)) || !std::isfinite(L.t0) || !std::isfinite(L.tf) ||
                !(ramp ? L.tf > L.t0 : L.tf >= L.t0))
                return h->fail(MPS_ERR_ARG, at + "surface load needs a finite value and window (tf > t0 for a ramp)");
            for (int k = 0; k < L.n; k++) {
                if (L.index[k] < 0 || L.index[k] >= Ne)
                    return h->fail(MPS_ERR_ARG, at + fmt("element %d outside [0,%lld)", L.index[k], (long long)Ne));
                if (L.face[k] < 1 || L.face[k] > (nen == 8 ? 6 : 4))
                    return h->fail(MPS_ERR_ARG, at + fmt("face %d outside 1..%d", L.face[k], nen == 8 ? 6 : 4));
            }
            n_face += L.n;
        } else {
            return h->fail(MPS_ERR_ARG, at + fmt("unknown type %d", L.type));
        }
    }
    const int per_face = (nen == 8) ? 4 : 1;
    const int64_t n_sgp = n_face * per_face;
    if (n_force >= (int64_t)MPS_ITEM_INDEX || n_sgp * 8 >= (int64_t)MPS_ITEM_INDEX)
        return h->fail(MPS_ERR_ARG, "mps_set_loads: too many load entries");
    (void)hipSetDevice(h->device);
    // reference geometry back from the device (the handle keeps no host copy)
    std::vector<double> X(3 * (size_t)N);
    std::vector<int32_t> conn((size_t)nen * Ne);
    std::vector<uint32_t> ninfo((size_t)N);
    HCHK(h, hipMemcpy(X.data(), I.X, X.size() * 8, hipMemcpyDeviceToHost));
    HCHK(h, hipMemcpy(conn.data(), I.conn, conn.size() * 4, hipMemcpyDeviceToHost));
    HCHK(h, hipMemcpy(ninfo.data(), I.ninfo, (size_t)N * 4, hipMemcpyDeviceToHost));
    for (uint32_t& f : ninfo) f &= MPS_REC_MASK;
    std::vector<LoadItem> items;
    std::vector<double> lm, lwin, sgd, sblk(MPS_SBLK_STRIDE * (size_t)std::max(n_loads, 1), 0.0);
    std::vector<int32_t> sge, sgb;
    int64_t seq = 0;
    for (int b = 0; b < n_loads; b++) {
        const mps_load& L = loads[b];
        double* B = sblk.data() + MPS_SBLK_STRIDE * (size_t)b;
        B[0] = (double)L.type;
        B[1] = L.value;
        B[2] = L.t0;
        B[3] = L.tf;
        if (L.type >= MPS_LOAD_TRACTION_RAMP)
            for (int d = 0; d < 3; d++) B[4 + d] = L.force[d];
        if (L.type == MPS_LOAD_FORCE) {
            for (int k = 0; k < L.n; k++) {
                const size_t j = lwin.size() / 2;
                for (int d = 0; d < 3; d++) lm.push_back(L.force[3 * (size_t)k + d] / (L.tf - L.t0));  // classNodes.cpp:93-95
                lwin.push_back(L.t0);
                lwin.push_back(L.tf);
                items.push_back({L.index[k], b, seq++, (MPS_ITEM_FORCE << 30) | (uint32_t)j});
            }
        } else if (L.type != MPS_LOAD_GRAVITY) {
            const bool ramp = L.type == MPS_LOAD_PRESSURE_RAMP;  // bulk-shape scatter (tractions use the face's)
            for (int k = 0; k < L.n; k++) {
                double pts[4 * MPS_SGP_STRIDE];
                int32_t nodes[32];
                const int32_t el = h->epos.empty() ? L.index[k] : h->epos[(size_t)L.index[k]];  // stored element
                const int ng = surface_points(X, N, conn.data() + (size_t)el * nen, nen, L.face[k], ramp, pts, nodes);
                for (int g = 0; g < ng; g++) {
                    const size_t q = sge.size();
                    sge.push_back(el);
                    sgb.push_back(b);
                    sgd.insert(sgd.end(), pts + (size_t)g * MPS_SGP_STRIDE, pts + (size_t)(g + 1) * MPS_SGP_STRIDE);
                    for (int a = 0; a < 8; a++)
                        if (nodes[g * 8 + a] >= 0)
                            items.push_back({nodes[g * 8 + a], b, seq++, (MPS_ITEM_PRESSURE << 30) | (uint32_t)(q * 8 + a)});
                }
            }
        }
    }
    if (grav_block >= 0) {  // gravity at its place among the other items of every loaded node
        std::vector<char> has((size_t)N, 0);
        for (const LoadItem& it : items) has[(size_t)it.node] = 1;
        for (int64_t n = 0; n < N; n++)
            if (has[(size_t)n]) items.push_back({(int32_t)n, grav_block, -1, MPS_ITEM_GRAVITY << 30});
    }
    std::sort(items.begin(), items.end(), [](const LoadItem& x, const LoadItem& y) {
        if (x.node != y.node) return x.node < y.node;
        if (x.block != y.block) return x.block < y.block;
        return x.seq < y.seq;
    });
    // flagged nodes in ascending order, each with its items (CSR)
    std::vector<int32_t> lidx((size_t)N, -1), lnn, lrow(1, 0), lcode(items.size());
    for (size_t j = 0; j < items.size(); j++) {
        const int32_t n = items[j].node;
        if (lnn.empty() || lnn.back() != n) {
            lidx[(size_t)n] = (int32_t)lnn.size();
            lnn.push_back(n);
            lrow.push_back(lrow.back());
            ninfo[(size_t)n] |= MPS_NODE_LOADED;
        }
        lcode[j] = (int32_t)items[j].code;
        lrow.back()++;
    }
    HCHK(h, hipStreamSynchronize(h->stream));
    for (void* q : h->load_allocs) (void)hipFree(q);
    h->load_allocs.clear();
    I.n_lnode = 0;
    I.n_sgp = 0;
    I.lidx = I.lnn = I.lrow = I.lcode = I.sge = I.sgb = nullptr;
    I.lm = I.lwin = I.fgrav = I.sgd = I.sblk = nullptr;
    I.pc = I.lf = I.lfprev = nullptr;
    auto upload = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
        hipError_t e = hipMalloc(dst, bytes < 256 ? 256 : bytes);
        if (e != hipSuccess) return e;
        h->load_allocs.push_back(*dst);
        return src ? hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    void* p;
    if (!items.empty()) {
        HCHK(h, upload(lidx.data(), lidx.size() * 4, &p));
        I.lidx = (const int32_t*)p;
        HCHK(h, upload(lnn.data(), lnn.size() * 4, &p));
        I.lnn = (const int32_t*)p;
        HCHK(h, upload(lrow.data(), lrow.size() * 4, &p));
        I.lrow = (const int32_t*)p;
        HCHK(h, upload(lcode.data(), lcode.size() * 4, &p));
        I.lcode = (const int32_t*)p;
        HCHK(h, upload(nullptr, lnn.size() * 24, &p));
        I.lf = (double*)p;
        I.n_lnode = (int32_t)lnn.size();
        if (I.energy_every > 0) {  // previous step's external force of the flagged nodes (work of the loads)
            HCHK(h, upload(nullptr, lnn.size() * 24, &p));
            HCHK(h, hipMemset(p, 0, lnn.size() * 24));
            I.lfprev = (double*)p;
        }
    }
    if (!lm.empty()) {
        HCHK(h, upload(lm.data(), lm.size() * 8, &p));
        I.lm = (const double*)p;
        HCHK(h, upload(lwin.data(), lwin.size() * 8, &p));
        I.lwin = (const double*)p;
    }
    if (!sge.empty()) {
        HCHK(h, upload(sge.data(), sge.size() * 4, &p));
        I.sge = (const int32_t*)p;
        HCHK(h, upload(sgb.data(), sgb.size() * 4, &p));
        I.sgb = (const int32_t*)p;
        HCHK(h, upload(sgd.data(), sgd.size() * 8, &p));
        I.sgd = (const double*)p;
        HCHK(h, upload(sblk.data(), sblk.size() * 8, &p));
        I.sblk = (const double*)p;
        HCHK(h, upload(nullptr, sge.size() * 24 * 8, &p));
        I.pc = (double*)p;
        I.n_sgp = (int32_t)sge.size();
    }
    HCHK(h, hipMemcpy((void*)I.ninfo, ninfo.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    if (grav_block >= 0 && grav != 0.0) {
        HCHK(h, upload(nullptr, (size_t)N * 8, &p));
        // the partial-force buffer (>= nen*Ne doubles) is free until the first step
        h->part_valid = false;
        hipError_t e = mpsk::launch_gravity(I, grav, I.part, (double*)p, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("gravity load: %s", hipGetErrorString(e)));
        I.fgrav = (const double*)p;
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

// mps_get_dofs staged through the device's fixed `scratch` array: no allocation in the call.  Nothing in the
// array needs to start at zero (the call writes the n indices it reads and the n values it copies out), and
// the call returns only after its copies finished, so calls on one device race only when they overlap.
extern "C" int mps_get_dofs_static(mps_handle h, int32_t which, int64_t n, const int32_t* dof, double* out) {
    if (!h || which < 0 || which > 2 || n < 0 || (n > 0 && (!dof || !out))) return MPS_ERR_ARG;
    if (n > MPS_DOFS_STATIC_MAX)
        return h->fail(MPS_ERR_CAPACITY, fmt("mps_get_dofs_static: %lld DOFs requested, at most %d per call",
                                             (long long)n, MPS_DOFS_STATIC_MAX));
    if (n == 0) return MPS_OK;
    const int64_t ndof = 3 * (int64_t)h->I.N;
    for (int64_t i = 0; i < n; i++)
        if (dof[i] < 0 || dof[i] >= ndof)
            return h->fail(MPS_ERR_ARG, fmt("mps_get_dofs_static: dof %d outside [0, %lld)", dof[i], (long long)ndof));
    (void)hipSetDevice(h->device);
    hipError_t why = hipSuccess;
    char* base = (char*)scratch_base(h->device, &why);
    if (!base)
        return h->fail(MPS_ERR_HIP, fmt("mps_get_dofs_static: device staging array not resolved: %s", hipGetErrorString(why)));
    int32_t* idx = reinterpret_cast<int32_t*>(base);
    double* vals = reinterpret_cast<double*>(base + MPS_STATIC_IDX_BYTES);
    const double* src = which == 0 ? h->I.u : (which == 1 ? h->I.v : h->I.a);
    HCHK(h, hipMemcpyAsync(idx, dof, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pick, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, src, idx, n, vals);
    HCHK(h, hipGetLastError());
    HCHK(h, hipMemcpyAsync(out, vals, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    return MPS_OK;
}

// Internal force at DOF idx[i] = node + N*dir: the node's partials in ascending patch order, summed as the node
// kernel sums them (k_gather_fint).
__global__ void k_pick_fint(MpsInst I, const int32_t* __restrict__ idx, int64_t n, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dof = idx[i], node = dof % I.N, dir = dof / I.N;
    double f = 0.0;
    for (int q = I.nrow[node]; q < I.nrow[node + 1]; q++) {
        double c[3];
        mpsd::load_partial(I, I.ncol[q], c);
        f += c[dir];
    }
    out[i] = f;
}

extern "C" int mps_get_internal_force_dofs(mps_handle h, int64_t n, const int32_t* dof, double* out) {
    if (!h || n < 0 || (n > 0 && (!dof || !out))) return MPS_ERR_ARG;
    if (n == 0) return MPS_OK;
    const int64_t ndof = 3 * (int64_t)h->I.N;
    for (int64_t i = 0; i < n; i++)
        if (dof[i] < 0 || dof[i] >= ndof)
            return h->fail(MPS_ERR_ARG, fmt("mps_get_internal_force_dofs: dof %d outside [0, %lld)", dof[i], (long long)ndof));
    (void)hipSetDevice(h->device);
    if (!h->part_valid) {  // before the first step: the forces of the current u (no viscous part: no increment yet)
        MpsInst I = h->I;
        I.ctl = nullptr;
        hipError_t e = mpsk::launch_elem(I, 1, h->stream);
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("internal force: %s", hipGetErrorString(e)));
        h->part_valid = true;
    }
    if (n > h->pick_cap) {
        HCHK(h, hipStreamSynchronize(h->stream));
        if (h->pick_idx) (void)hipFree(h->pick_idx);
        if (h->pick_out) (void)hipFree(h->pick_out);
        h->pick_idx = nullptr;
        h->pick_out = nullptr;
        h->pick_cap = 0;
        HCHK(h, hipMalloc(&h->pick_idx, (size_t)n * 4));
        HCHK(h, hipMalloc(&h->pick_out, (size_t)n * 8));
        h->pick_cap = n;
    }
    HCHK(h, hipMemcpyAsync(h->pick_idx, dof, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pick_fint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->I, h->pick_idx, n,
                       h->pick_out);
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

// Launch sequence of `count` steps starting at h->next_step; the first step is begun by the standalone begin
// kernel unless it already is; the last node pass is unfused.  With energies every `every` steps, step k
// computes DEFO_ENERGY and VOLUME of u_k (k % every == 0) and KIN_ENERGY of step k-1 from the velocity that
// still holds it before the node pass (when (k-1) % every == 0 and k-1 is in this chunk), in the element kernel
// (FP64 inviscid form) or in the standalone pass before the node kernel; after the last node pass, the pass for
// the KIN_ENERGY of the chunk's last step.  The node kernels accumulate the work sums every step.
static hipError_t enqueue_steps(mps_solver_s* h, long long count, std::vector<int>* kinds) {
    const MpsInst& I = h->I;
    const int every = h->cfg.energy_every;
    const bool fused = every > 0 && mpsk::elem_fuses_energy(I);
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
        unsigned ef = 0u;
        if (every > 0 && k % every == 0) ef |= MPS_E_DEF;
        if (every > 0 && i >= 1 && k >= 2 && (k - 1) % every == 0) ef |= MPS_E_KIN;
        {
            TimedLaunch t(h, 0);
            e = mpsk::launch_elem(I, 0, h->stream, fused ? ef : 0u);
            t.done();
            if (t.a) kinds->push_back(0);
        }
        if (e == hipSuccess && ef && !fused) {
            TimedLaunch t(h, 3);
            e = mpsk::launch_energy(I, ef, h->stream);
            t.done();
            if (t.a) kinds->push_back(3);
        }
        // between two node kernels of a launch the state holds the half-step velocity w instead of v and a: in
        // v's array, or in a's after a step whose v the next element phase reads (kinetic energy); the first
        // node kernel of a launch finds the v and a that k_begin leaves in place
        const bool out_v = every > 0 && k % every == 0;
        const unsigned vaf = (i == 0 ? MPS_N_IN_VA : (every > 0 && (k - 1) % every == 0) ? MPS_N_IN_WA : 0u) |
                             (out_v ? MPS_N_OUT_V : 0u);
        if (e == hipSuccess && (I.n_sgp > 0 || I.n_lnode > 0)) e = mpsk::launch_loads(I, h->stream);
        if (e == hipSuccess && I.n_series > 0) e = mpsk::launch_series(I, h->stream, vaf);
        if (e != hipSuccess) break;
        const bool last = (i == count - 1);
        {
            TimedLaunch t(h, 1);
            e = mpsk::launch_node(I, last ? 0 : 1, h->stream, ef | vaf | (i >= 1 ? MPS_N_PREV : 0u));
            t.done();
            if (t.a) kinds->push_back(1);
        }
        if (e != hipSuccess) break;
        if (last && every > 0) {  // the kernels check that the last step that ran is an energy step
            TimedLaunch t(h, 3);
            e = mpsk::launch_energy(I, MPS_E_KIN_NOW, h->stream);
            if (e == hipSuccess) e = mpsk::launch_energy_tail(I, h->stream);
            t.done();
            if (t.a) kinds->push_back(3);
        }
        h->next_step = k + 1;
        h->step_begun = !last;
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

// Chunks of at most kChunk steps: the host reads the control block back once per chunk.
static const long long kChunk = 256;

// True when mps_run replays its chunks from captured graphs in the handle's current state.
static bool graph_replay(const mps_solver_s* h) { return h->use_graphs && !h->profiling && !h->step_begun; }

// The graph that replays a chunk of n steps (begin, (element, node+begin) x (n-1), element, node), captured
// and instantiated on first use.  Capture does not execute anything: the host mirror of the step counter is
// restored afterwards.  The kernels read the step index from the control block, so one graph serves every
// chunk of n steps wherever it starts.
static hipError_t chunk_graph(mps_solver_s* h, long long n, hipGraphExec_t* out) {
    const long long phase = h->cfg.energy_every > 0 ? h->next_step % h->cfg.energy_every : 0;
    for (auto& g : h->graphs)
        if (g.n == n && g.phase == phase) {
            *out = g.exec;
            return hipSuccess;
        }
    std::vector<int> kinds;
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    const long long k0 = h->next_step;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = enqueue_steps(h, n, &kinds);
    hipError_t e2 = hipStreamEndCapture(h->stream, &gr);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    h->next_step = k0;
    h->step_begun = false;
    if (e != hipSuccess) return e;
    h->graphs.push_back({n, phase, ge});
    *out = ge;
    return hipSuccess;
}

// Chunk lengths mps_run(h, max_steps) would use from the current step.
static std::vector<long long> chunk_plan(const mps_solver_s* h, long long max_steps) {
    long long remaining = max_steps;
    if (h->I.time_mode == MPS_TIME_FIXED) remaining = std::min<long long>(remaining, h->I.n_steps - (h->next_step - 1));
    std::vector<long long> plan;
    while (remaining > 0) {
        const long long n = std::min(remaining, kChunk);
        plan.push_back(n);
        remaining -= n;
    }
    return plan;
}

extern "C" int mps_prepare_run(mps_handle h, int64_t max_steps) {
    if (!h) return MPS_ERR_ARG;
    if (max_steps < 0) return h->fail(MPS_ERR_ARG, "mps_prepare_run: max_steps < 0");
    (void)hipSetDevice(h->device);
    if (!graph_replay(h)) return MPS_OK;  // mps_run will launch the kernels one by one
    const long long k0 = h->next_step;
    for (long long n : chunk_plan(h, max_steps)) {  // each chunk's graph depends on its first step (energies)
        hipGraphExec_t ge = nullptr;
        hipError_t e = chunk_graph(h, n, &ge);
        if (e != hipSuccess) {
            h->next_step = k0;
            return h->fail(MPS_ERR_HIP, fmt("graph capture: %s", hipGetErrorString(e)));
        }
        h->next_step += n;
    }
    h->next_step = k0;
    return MPS_OK;
}

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
    float total_ms = 0.f;
    for (long long n : chunk_plan(h, max_steps)) {
        std::vector<int> kinds;
        hipError_t e = hipSuccess;
        if (graph_replay(h)) {
            // every chunk of n steps is the same launch sequence: replay it from a captured graph
            hipGraphExec_t ge = nullptr;
            e = chunk_graph(h, n, &ge);
            if (e == hipSuccess) e = hipGraphLaunch(ge, h->stream);
            if (e == hipSuccess) {
                h->next_step += n;
                h->step_begun = false;
            }
        } else {
            e = enqueue_steps(h, n, &kinds);
        }
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("step launch: %s", hipGetErrorString(e)));
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
    h->part_valid = true;  // the partial forces are those of the last completed step's u
    if (rep) *rep = r;
    return status_code(r.status);
}

// Output snapshots: the state of the last completed step staged on the device in stream order (so later
// steps may run at once), then copied to pinned host memory on a second stream.
extern "C" int mps_snapshot_begin(mps_handle h, int32_t slot, int32_t with_elements) {
    if (!h) return MPS_ERR_ARG;
    if (slot < 0 || slot > 1) return h->fail(MPS_ERR_ARG, "mps_snapshot_begin: slot must be 0 or 1");
    auto& S = h->snap[slot];
    if (S.busy.load()) return h->fail(MPS_ERR_STATE, "mps_snapshot_begin: slot not released (mps_snapshot_release)");
    (void)hipSetDevice(h->device);
    const size_t n3 = 3 * (size_t)h->I.N, ne = 21 * (size_t)h->I.Ne;
    const size_t cap = (3 * n3 + ne) * 8;
    if (!S.dev) {
        HCHK(h, hipMalloc(&S.dev, cap));
        HCHK(h, hipHostMalloc((void**)&S.host, cap, hipHostMallocDefault));
        HCHK(h, hipEventCreateWithFlags(&S.staged, hipEventDisableTiming));
        HCHK(h, hipEventCreateWithFlags(&S.copied, hipEventDisableTiming));
    }
    if (!h->copy_stream) HCHK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    const MpsInst& I = h->I;
    HCHK(h, hipMemcpyAsync(S.dev, I.u, n3 * 8, hipMemcpyDeviceToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(S.dev + n3, I.v, n3 * 8, hipMemcpyDeviceToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(S.dev + 2 * n3, I.a, n3 * 8, hipMemcpyDeviceToDevice, h->stream));
    if (with_elements) {
        hipError_t e = mpsk::launch_elem_outputs(I, S.dev + 3 * n3, h->stream);
        if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("element outputs: %s", hipGetErrorString(e)));
    }
    HCHK(h, hipEventRecord(S.staged, h->stream));
    HCHK(h, hipStreamWaitEvent(h->copy_stream, S.staged, 0));
    HCHK(h, hipMemcpyAsync(S.host, S.dev, (3 * n3 + (with_elements ? ne : 0)) * 8, hipMemcpyDeviceToHost,
                           h->copy_stream));
    HCHK(h, hipEventRecord(S.copied, h->copy_stream));
    S.with_elem = with_elements ? 1 : 0;
    S.busy.store(1);
    return MPS_OK;
}

// May run on another thread than the one that steps the solver: it touches only the slot.
extern "C" int mps_snapshot_wait(mps_handle h, int32_t slot, const double** u, const double** v, const double** a,
                                 const double** elem) {
    if (!h || slot < 0 || slot > 1) return MPS_ERR_ARG;
    auto& S = h->snap[slot];
    if (!S.busy.load()) return MPS_ERR_STATE;
    if (hipEventSynchronize(S.copied) != hipSuccess) return MPS_ERR_HIP;
    const size_t n3 = 3 * (size_t)h->I.N;
    if (u) *u = S.host;
    if (v) *v = S.host + n3;
    if (a) *a = S.host + 2 * n3;
    if (elem) *elem = S.with_elem ? S.host + 3 * n3 : nullptr;
    return MPS_OK;
}

extern "C" int mps_snapshot_release(mps_handle h, int32_t slot) {
    if (!h || slot < 0 || slot > 1) return MPS_ERR_ARG;
    h->snap[slot].busy.store(0);
    return MPS_OK;
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

// Columns col0 .. col0+ncol-1 of the history ring, entries [first, first+count), into outs[j] (NULL: skipped).
static int read_history(mps_solver_s* h, int64_t first, int64_t count, int col0, int ncol, double** outs) {
    const int cap = h->I.hist_cap;
    std::vector<double> buf((size_t)ncol * cap);
    HCHK(h, hipMemcpy(buf.data(), h->I.hist + (size_t)col0 * cap, buf.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < count; i++) {
        const size_t idx = (size_t)((first + i) % cap);
        for (int j = 0; j < ncol; j++)
            if (outs[j]) outs[j][i] = buf[(size_t)j * cap + idx];
    }
    return MPS_OK;
}

extern "C" int mps_get_history(mps_handle h, int64_t first, int64_t count, double* time, double* dt, double* e_def,
                               double* e_kin, double* volume) {
    if (!h) return MPS_ERR_ARG;
    if (first < 0 || count < 0) return h->fail(MPS_ERR_ARG, "mps_get_history: negative range");
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    double* outs[5] = {time, dt, e_def, e_kin, volume};
    return read_history(h, first, count, MPS_H_TIME, 5, outs);
}

extern "C" int mps_set_dof_series(mps_handle h, int64_t n, const int32_t* which, const int32_t* dof) {
    if (!h || n < 0 || (n > 0 && (!which || !dof))) return MPS_ERR_ARG;
    const int64_t ndof = 3 * (int64_t)h->I.N;
    for (int64_t i = 0; i < n; i++) {
        if (dof[i] < 0 || dof[i] >= ndof)
            return h->fail(MPS_ERR_ARG, fmt("mps_set_dof_series: dof %d outside [0, %lld)", dof[i], (long long)ndof));
        if (which[i] < MPS_SERIES_U || which[i] > MPS_SERIES_FEXT)
            return h->fail(MPS_ERR_ARG, fmt("mps_set_dof_series: unknown kind %d", which[i]));
    }
    if (n > INT32_MAX) return h->fail(MPS_ERR_ARG, "mps_set_dof_series: too many entries");
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);  // the captured steps launch the series kernel
    h->graphs.clear();
    for (void* p : h->series_allocs) (void)hipFree(p);
    h->series_allocs.clear();
    MpsInst& I = h->I;
    I.n_series = 0;
    I.ser_dof = I.ser_kind = nullptr;
    I.ser_ring = nullptr;
    if (n == 0) return MPS_OK;
    void *pd = nullptr, *pk = nullptr, *pr = nullptr;
    const size_t ring = (size_t)n * (size_t)I.hist_cap * 8;
    HCHK(h, hipMalloc(&pd, (size_t)n * 4));
    h->series_allocs.push_back(pd);
    HCHK(h, hipMalloc(&pk, (size_t)n * 4));
    h->series_allocs.push_back(pk);
    HCHK(h, hipMalloc(&pr, ring));
    h->series_allocs.push_back(pr);
    HCHK(h, hipMemcpy(pd, dof, (size_t)n * 4, hipMemcpyHostToDevice));
    HCHK(h, hipMemcpy(pk, which, (size_t)n * 4, hipMemcpyHostToDevice));
    hipError_t e = mpsk::launch_fill_u64((unsigned long long*)pr, (long long)n * I.hist_cap, 0x7FF8000000000000ull,
                                         h->stream);  // NaN until a step writes the row
    if (e != hipSuccess) return h->fail(MPS_ERR_HIP, fmt("series ring: %s", hipGetErrorString(e)));
    HCHK(h, hipStreamSynchronize(h->stream));
    I.n_series = (int32_t)n;
    I.ser_dof = (const int32_t*)pd;
    I.ser_kind = (const int32_t*)pk;
    I.ser_ring = (double*)pr;
    return MPS_OK;
}

extern "C" int mps_get_dof_series(mps_handle h, int64_t first, int64_t count, double* out) {
    if (!h || first < 0 || count < 0 || (count > 0 && !out)) return MPS_ERR_ARG;
    const MpsInst& I = h->I;
    if (I.n_series == 0 || count == 0) return MPS_OK;
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    const size_t n = (size_t)I.n_series, cap = (size_t)I.hist_cap;
    std::vector<double> buf(n * cap);
    HCHK(h, hipMemcpy(buf.data(), I.ser_ring, buf.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < count; r++)
        std::memcpy(out + (size_t)r * n, buf.data() + (size_t)((first + r) % (int64_t)cap) * n, n * 8);
    return MPS_OK;
}

extern "C" int mps_get_energy_balance(mps_handle h, int64_t first, int64_t count, double* e_ext, double* e_kin_lumped,
                                      double* w_ext, double* w_react) {
    if (!h) return MPS_ERR_ARG;
    if (first < 0 || count < 0) return h->fail(MPS_ERR_ARG, "mps_get_energy_balance: negative range");
    (void)hipSetDevice(h->device);
    HCHK(h, hipStreamSynchronize(h->stream));
    double* outs[4] = {e_ext, e_kin_lumped, w_ext, w_react};
    return read_history(h, first, count, MPS_H_EEXT, 4, outs);
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
    // its own partial forces, so that the step's stay valid for mps_get_internal_force_dofs
    if (!h->fint_part) HCHK(h, dalloc(h, &h->fint_part, std::max<size_t>((size_t)h->I.n_part * 3 * 8, 256)));
    MpsInst I = h->I;
    I.u = h->scratch_u;
    I.ctl = nullptr;
    I.part = h->fint_part;
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

extern "C" int mps_layout(mps_handle h, int64_t* n_patch, int64_t* n_part, int32_t* elem_order) {
    if (!h) return MPS_ERR_ARG;
    if (n_patch) *n_patch = h->I.n_patch;
    if (n_part) *n_part = h->I.n_part;
    if (elem_order) {
        const size_t Ne = (size_t)h->I.Ne;
        if (h->epos.empty())
            for (size_t s = 0; s < Ne; s++) elem_order[s] = (int32_t)s;
        else
            for (size_t e = 0; e < Ne; e++) elem_order[(size_t)h->epos[e]] = (int32_t)e;
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
