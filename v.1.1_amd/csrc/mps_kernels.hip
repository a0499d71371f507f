// mps_kernels.hip -- HIP kernels of the explicit step (gfx950).
//
// Per step (fixed pattern, no host round trip):
//   element kernel  : u_k -> element nodal forces (T = double or float), summed per patch in
//                     LDS into FP64 partial nodal forces, and the element-wise stable dt for
//                     step k+1 (block min, 64-bit integer atomicMin)
//   node kernel     : CSR gather of the partials (ascending patch order, FP64, no float atomics),
//                     a = -fint/m, v += dt(a_old + a_new)/2, and -- when fused -- the
//                     begin of step k+1 (lagged time, dt, prediction, Dirichlet BC)
// plus a standalone begin kernel (first step of a launch) and an energy pass.
#include "../../include/mps_gpu.h"
#include "mps_device.h"
#include "mps_kernels.h"

using namespace mpsd;

#define INF_BITS 0x7FF0000000000000ull

// ---------------------------------------------------------------------------
// Element kernels
// ---------------------------------------------------------------------------
// The patch's lists for the kernels with one work-item per element: its contribution list (pcol,
// NEN*ne entries, 16-byte aligned) and its node offsets (prow, U+1 entries from entry poff[p]+p,
// staged from the 8-entry boundary below).  Each work-item loads its 8-entry pieces (pc8: entries
// q..q+7 when q < m; the single-instance set-up pads both arrays by 8 entries, so a piece that
// straddles m stays inside the allocation and the entries past m are never used) into registers at
// the start of the patch, so that their latency hides behind the element work, and writes them to LDS
// just before the barrier that ends the element phase.
__device__ __forceinline__ uint4 pc8(const uint16_t* __restrict__ g, int q, int m) {
    return q < m ? *reinterpret_cast<const uint4*>(g + q) : make_uint4(0u, 0u, 0u, 0u);
}

// ROWS = false: only the contribution list; the node offsets are then read from prow in global memory.
template <int NEN, bool ROWS = true>
struct PatchList {
    static constexpr int NQ = (NEN * MPS_PATCH + 8 * MPS_PATCH - 1) / (8 * MPS_PATCH);  // pcol pieces per work-item
    static constexpr int NR = ROWS ? ((NEN * MPS_PATCH + 1 + 7 + 7) / 8 + MPS_PATCH - 1) / MPS_PATCH : 0;  // prow pieces
    static constexpr int SPR = (NEN * MPS_PATCH + 15) / 8 * 8;  // LDS entries for the staged prow (>= len rounded up to 8)
    uint4 v[NQ], w[NR > 0 ? NR : 1];
    int off, len;  // first prow entry of the patch relative to the staged start; staged entries
    __device__ __forceinline__ void load(const MpsInst& I, int p, int e0, int m, int U) {
        const uint16_t* gpc = I.pcol + (size_t)e0 * NEN;
#pragma unroll
        for (int r = 0; r < NQ; r++) v[r] = pc8(gpc, (threadIdx.x + r * MPS_PATCH) * 8, m);
        const int s = I.poff[p] + p, a0 = s & ~7;
        off = s - a0;
        len = off + U + 1;
#pragma unroll
        for (int r = 0; r < NR; r++) w[r] = pc8(I.prow + a0, (threadIdx.x + r * MPS_PATCH) * 8, len);
    }
    __device__ __forceinline__ void store(uint16_t* spc, uint16_t* spr, int m) const {
#pragma unroll
        for (int r = 0; r < NQ; r++) {
            const int q = (threadIdx.x + r * MPS_PATCH) * 8;
            if (q < m) *reinterpret_cast<uint4*>(spc + q) = v[r];
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int q = (threadIdx.x + r * MPS_PATCH) * 8;
            if (q < len) *reinterpret_cast<uint4*>(spr + q) = w[r];
        }
    }
};

// Second half of a patch: after the barrier that follows the LDS force writes, the block sums, per
// distinct node of the patch, its contributions in ascending (element, slot) order in FP64 and writes
// one partial force per node.  With PRE (one work-item per element, BS = MPS_PATCH) the contribution
// list (spc) and the node offsets (spr, the patch's first at spr[0]) are in LDS; otherwise (the
// persistent batched kernels) both are read from global memory as they are needed.
template <typename T>
__device__ __forceinline__ void node_sum(const MpsInst& I, int g, int q0, int q1, int nen, const T* __restrict__ sf,
                                         const uint16_t* __restrict__ pc) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int q = q0; q < q1; q++) {
        const int c = pc[q];
        const int i = c / nen, a = c % nen;
        s0 += (double)sf[(a * 3 + 0) * MPS_PATCH + i];
        s1 += (double)sf[(a * 3 + 1) * MPS_PATCH + i];
        s2 += (double)sf[(a * 3 + 2) * MPS_PATCH + i];
    }
    if (sizeof(T) == 4) {  // precision form: partials rounded once to FP32, summed in FP64 by the node kernel
        float* dst = reinterpret_cast<float*>(I.part) + 3 * (size_t)g;
        dst[0] = (float)s0;
        dst[1] = (float)s1;
        dst[2] = (float)s2;
    } else {
        double* dst = I.part + 3 * (size_t)g;
        dst[0] = s0;
        dst[1] = s1;
        dst[2] = s2;
    }
}

template <typename T, int NEN, int BS, bool PRE>
__device__ __forceinline__ void patch_sums(const MpsInst& I, int p, const T* __restrict__ sf,
                                           const uint16_t* __restrict__ spc, const uint16_t* __restrict__ spr) {
    static_assert(!PRE || BS == MPS_PATCH, "PRE needs one work-item per element");
    const int base = I.poff[p];
    const int U = I.poff[p + 1] - base;
    if constexpr (PRE) {
        for (int j = threadIdx.x; j < U; j += BS) node_sum<T>(I, base + j, spr[j], spr[j + 1], NEN, sf, spc);
    } else {
        const uint16_t* __restrict__ pr = I.prow + base + p;
        const uint16_t* __restrict__ pc = I.pcol + (size_t)p * MPS_PATCH * NEN;
        for (int j = threadIdx.x; j < U; j += BS) node_sum<T>(I, base + j, pr[j], pr[j + 1], NEN, sf, pc);
    }
}

// Forces of one patch (elements [p*MPS_PATCH, ...)): each element's nodal forces go to LDS
// (SoA sf[(a*3+d)*MPS_PATCH + i]); after a barrier the block sums, per distinct node of the
// patch, its contributions in ascending (element, slot) order in FP64 and writes one partial
// force per node to part.  No float atomics: the node kernel adds the partials of a node in
// ascending patch order, so the sum order is fixed.  Returns the block's min element dt
// candidate (64-bit pattern, lanes other than those computing elements hold +inf).
// EN: the elements' energies (hex8_element / tet4_element) are added to env[3]; hex8 keeps its GPs' wJ in
// swj[gp*MPS_PATCH + i].
template <typename T, int NEN, int LAW, int BS, bool VISC = false, bool PRE = false, bool EN = false,
          bool O32 = false>
__device__ __forceinline__ unsigned long long patch_forces(const MpsInst& I, int p, bool want_dt, T* __restrict__ sf,
                                                           uint16_t* __restrict__ spc, uint16_t* __restrict__ spr,
                                                           double rdt = 0.0, unsigned eflags = 0u,
                                                           double* env = nullptr, double* swj = nullptr,
                                                           double dtsk = __builtin_inf()) {
    const int e0 = p * MPS_PATCH;
    const int ne = min(MPS_PATCH, I.Ne - e0);
    PatchList<NEN> pl;
    if constexpr (PRE) pl.load(I, p, e0, ne * NEN, I.poff[p + 1] - I.poff[p]);
    unsigned long long dmin = INF_BITS;
    for (int i = threadIdx.x; i < ne; i += BS) {
        T f[NEN][3];
        double en[3] = {0.0, 0.0, 0.0};
        const double dte =
            (NEN == 8) ? hex8_element<T, LAW, VISC, EN, O32>(I, e0 + i, I.u, (T(*)[3])f, want_dt,
                                                              EN ? (T*)(swj + i) : sf + i, rdt, eflags, en, dtsk)
                       : tet4_element<T, LAW, VISC, EN, O32>(I, e0 + i, I.u, (T(*)[3])f, want_dt, rdt, eflags, en);
        if constexpr (EN) {
#pragma unroll
            for (int j = 0; j < 3; j++) env[j] += en[j];
        }
#pragma unroll
        for (int a = 0; a < NEN; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) sf[(a * 3 + d) * MPS_PATCH + i] = f[a][d];
        dmin = umin64(dmin, d2bits(dte));
    }
    if constexpr (PRE) pl.store(spc, spr, ne * NEN);
    __syncthreads();
    patch_sums<T, NEN, BS, PRE>(I, p, sf, spc, spr + (PRE ? pl.off : 0));
    return dmin;
}

// Energies of the step in flight from the element phase's per-work-item sums env (DEFO_ENERGY, VOLUME of step k,
// KIN_ENERGY of step k-1): the block's sums go to the element region of epart, which the node kernel of step k
// adds up.  sf (the force slots, no longer needed once the patch sums are done) is the block's scratch.
__device__ __forceinline__ void elem_energy_publish(const MpsInst& I, double* env, double* sf) {
    __syncthreads();
    publish_block_sum<3>(env, I.epart, sf);
}

// Exact stable dt of element e at u (the element kernels' own arithmetic, bit for bit).
template <int NEN, bool O32>
__device__ __forceinline__ double elem_dt_exact(const MpsInst& I, int e) {
    int nd[NEN];
#pragma unroll
    for (int a = 0; a < NEN; a++) nd[a] = I.conn[(size_t)e * NEN + a];
    double pa, pb, c;
    elem_mat(I, e, pa, pb, c);
    double xX[NEN][3], xU[NEN][3], x[NEN][3];
    gather_nodes<NEN, O32, true>((size_t)I.N, nd, I.X, xX, I.u, xU);
#pragma unroll
    for (int a = 0; a < NEN; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) x[a][d] = xX[a][d] + xU[a][d];
    return elem_dt(NEN == 8 ? hex_char_length(x) : tet_char_length(x), I.l0[e], c);
}

// End of an FP64 element kernel of step k that may have skipped elements (dt threshold dtsk).  The last
// workgroup to finish checks the certificate: the step's min over the exactly computed elements is at most
// dtsk, so no skipped element (each certainly above dtsk) can be below it.  Otherwise this workgroup adds every
// element's exact dt to the min and skipping stops for MPS_DT_SKIP_HOLD steps.  Either way the min the node
// kernel reads is the min over all elements, bit for bit.  The count is two-level, so that thousands of
// workgroups do not queue on one address (one counter for all made the tet4 kernel twice as slow): the last
// workgroup of each residue modulo MPS_DT_SLOTS counts its residue on the top counter, and each last one
// resets the counter it completed.  No fence orders a workgroup's atomicMin before its
// count (a fence per workgroup made the kernel 80% slower), so the certifying workgroup may read a min that
// misses some workgroups' values: one not yet lowered by them.  That is safe both ways: a min at most dtsk
// only falls further, and one above it only sends the step through the exact fallback, which the certifying
// workgroup therefore takes only after reading the slots again over up to 8 pauses of about 8K cycles.
template <int NEN, bool O32>
__device__ void dt_certify(const MpsInst& I, long long k, double dtsk, unsigned long long* sh) {
    __shared__ int last, fail;
    if (threadIdx.x == 0) {
        const unsigned r = blockIdx.x & (MPS_DT_SLOTS - 1);
        const unsigned share = (gridDim.x - r + MPS_DT_SLOTS - 1) / MPS_DT_SLOTS;  // workgroups of residue r
        const unsigned groups = gridDim.x < MPS_DT_SLOTS ? gridDim.x : MPS_DT_SLOTS;
        last = 0;
        if (atomicAdd(dt_count(I, r), 1u) == share - 1) {
            *dt_count(I, r) = 0u;
            last = atomicAdd(dt_count(I, MPS_DT_SLOTS), 1u) == groups - 1;
            if (last) *dt_count(I, MPS_DT_SLOTS) = 0u;
        }
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x < 64) {  // wave 0 decides for the workgroup
        unsigned long long m = dt_read_min_now(I, (k + 1) & 1);
        for (int t = 0; t < 8 && !(bits2d(m) <= dtsk); t++) {
            __builtin_amdgcn_s_sleep(127);
            m = dt_read_min_now(I, (k + 1) & 1);
        }
        if (threadIdx.x == 0) fail = !(bits2d(m) <= dtsk);
    }
    __syncthreads();
    if (fail) {
        unsigned long long mm = INF_BITS;
        for (int e = threadIdx.x; e < I.Ne; e += blockDim.x) mm = umin64(mm, d2bits(elem_dt_exact<NEN, O32>(I, e)));
        mm = block_min_u64(mm, sh);
        if (threadIdx.x == 0) {
            atomicMin(dt_slot(I, (k + 1) & 1, 0), mm);
            I.ctl->skip_hold = k + 1 + MPS_DT_SKIP_HOLD;
        }
    }
}

// mode 0: step (reads ctl->A, writes ctl->B, reduces dt for step k+1)
// mode 1: standalone force evaluation (no ctl access)
// One workgroup per patch, one element per work-item.
// VISC (mode 0 only): bulk artificial viscosity over the step, from u - uprev and the step's dt.
// EN (mode 0, FP64 inviscid): the energies of eflags (MPS_E_DEF, MPS_E_KIN) into the history.
// O32: node gathers by 32-bit offsets (gather_nodes; meshes of at most MPS_OFF32_NODES nodes).
// FP64 inviscid hex8 (SKIP): elements certainly above the step's dt threshold skip the strict length
// (hex8_element, dt_certify).
template <typename T, int NEN, int LAW, bool VISC = false, bool EN = false, bool O32 = false>
__global__ __launch_bounds__(MPS_PATCH) __attribute__((amdgpu_waves_per_eu(VISC && LAW == 1 ? 1 : 2))) void k_elem(MpsInst I, int mode, unsigned eflags) {
    __shared__ T sf[NEN * 3 * MPS_PATCH];
    __shared__ __align__(16) uint16_t spc[NEN * MPS_PATCH];
    __shared__ __align__(16) uint16_t spr[PatchList<NEN>::SPR];
    __shared__ unsigned long long sh[MPS_PATCH / 64];
    __shared__ double swj[(EN && NEN == 8) ? 8 * MPS_PATCH : 1];  // hex8 GP weights wJ for the kinetic energy
    constexpr bool SKIP = sizeof(T) == 8 && NEN == 8 && !VISC;
    long long k = 0;
    double rdt = 0.0, dtsk = __builtin_inf();
    if (mode == 0) {
        k = I.ctl->A;
        if (blockIdx.x == 0 && threadIdx.x == 0) I.ctl->B = k;
        const MpsStepSlot s = I.ctl->slot[k & 1];
        if (!s.go) return;
        if (VISC) rdt = (s.dt == 0.0) ? 0.0 : 1 / s.dt;
        if (SKIP) dtsk = I.ctl->dt_skip;
    }
    const int p = xcd_tile(blockIdx.x, gridDim.x);
    double env[3] = {0.0, 0.0, 0.0};
    unsigned long long m = patch_forces<T, NEN, LAW, MPS_PATCH, VISC, true, EN, O32>(I, p, mode == 0, sf, spc, spr,
                                                                                    rdt, eflags, env, swj, dtsk);
    if (mode == 0) {
        m = block_min_u64(m, sh);
        if (threadIdx.x == 0) atomicMin(dt_slot(I, (k + 1) & 1, blockIdx.x & (MPS_DT_SLOTS - 1)), m);
        if constexpr (EN) {
            // hex8 kinetic energy of the current v, after the force phase and its barriers: kept apart from the
            // force arithmetic, whose compiled form (and rounding) it would otherwise share a basic block with
            if (NEN == 8 && (eflags & MPS_E_KIN)) {
                const int e0 = p * MPS_PATCH;
                if ((int)threadIdx.x < min(MPS_PATCH, I.Ne - e0)) {
                    const int e = e0 + threadIdx.x;
                    const int4 q0 = *reinterpret_cast<const int4*>(I.conn + (size_t)e * 8);
                    const int4 q1 = *reinterpret_cast<const int4*>(I.conn + (size_t)e * 8 + 4);
                    const int nd[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
                    double wJ[8];
#pragma unroll
                    for (int gp = 0; gp < 8; gp++) wJ[gp] = swj[gp * MPS_PATCH + threadIdx.x];
                    env[2] += hex_kinetic<O32>(I, nd, I.rho ? I.rho[e] : I.rho0, wJ);
                }
            }
            elem_energy_publish(I, env, reinterpret_cast<double*>(sf));
        }
        if constexpr (SKIP) dt_certify<8, O32>(I, k, dtsk, sh);
    }
}

// Bytes of the tet4 kernel's LDS space shared by the node staging (48 per node) and the force slots
// (12*MPS_PATCH values of tsize bytes); a multiple of 16.
__host__ __device__ __forceinline__ size_t tet_shared_bytes(int pu, int tsize) {
    const size_t a = 48 * (size_t)pu, b = 12 * (size_t)MPS_PATCH * tsize;
    return a > b ? a : b;
}

// A staged node of the tet4 kernel: X and u of the patch's j-th distinct node at sxu[6j..6j+5], read as three
// 16-byte pieces (one address per node, the pieces at fixed offsets from it).
__device__ __forceinline__ void staged_node(const double* __restrict__ sxu, int j, double* x, double* u) {
    const double2* s = reinterpret_cast<const double2*>(sxu + 6 * j);
    const double2 a = s[0], b = s[1], c = s[2];
    x[0] = a.x;
    x[1] = a.y;
    x[2] = b.x;
    u[0] = b.y;
    u[1] = c.x;
    u[2] = c.y;
}

// Tet4 element kernel with node staging (inviscid).  Each block first loads the U_p distinct nodes of
// its patch (ascending ids, so the loads run along node rows) into LDS, 48 bytes per node; each
// element then reads its four nodes from LDS through its local indices plx.  Every node is fetched
// once per patch instead of once per element slot (about 1 against 5.8 node loads per partial on a
// structured tet mesh).  Dynamic LDS: the staging (48*pu_max bytes) and, after a barrier, the 12*MPS_PATCH
// force slots of T in the same space, then the 4*MPS_PATCH entries of the contribution list; the elements
// hold their forces in registers across that barrier, and the patch sums read the node offsets from global
// memory.  So 6 (FP64) or 8 (FP32) workgroups reside per CU.
template <typename T, int LAW, bool EN = false>
__global__ __launch_bounds__(MPS_PATCH) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 4))) void k_elem_tet_staged(MpsInst I, int mode, unsigned eflags) {
    extern __shared__ __align__(16) double smem[];  // 16-byte pieces are stored at offsets that are multiples of 16
    __shared__ unsigned long long sh[MPS_PATCH / 64];
    const int pu = I.pu_max;
    double* __restrict__ sxu = smem;           // [pu][6]: X (3), then u (3) of each staged node
    T* __restrict__ sf = reinterpret_cast<T*>(smem);  // [12][MPS_PATCH], over the staging once it is read
    const size_t shared_bytes = tet_shared_bytes(pu, sizeof(T));
    uint16_t* __restrict__ spc = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(smem) + shared_bytes);
    constexpr bool SKIP = sizeof(T) == 8;  // stable-dt skip (dt_certify)
    long long k = 0;
    double dtsk = __builtin_inf();
    if (mode == 0) {
        k = I.ctl->A;
        if (blockIdx.x == 0 && threadIdx.x == 0) I.ctl->B = k;
        const MpsStepSlot s = I.ctl->slot[k & 1];
        if (!s.go) return;
        if (SKIP) dtsk = I.ctl->dt_skip;
    }
    const int p = xcd_tile(blockIdx.x, gridDim.x);
    const int e0 = p * MPS_PATCH;
    const int ne = min(MPS_PATCH, I.Ne - e0);
    const int base = I.poff[p];
    const int U = I.poff[p + 1] - base;  // <= pu_max (set-up)
    const size_t N = (size_t)I.N;
    PatchList<4, false> pl;  // the node offsets stay in global memory (6 workgroups per CU in FP64)
    pl.load(I, p, e0, ne * 4, U);
    // the element's own inputs are loaded before the staging barrier, so their latency overlaps it
    const int i = threadIdx.x;
    const int e = e0 + min(i, ne - 1);
    const uint2 l4 = *reinterpret_cast<const uint2*>(I.plx + (size_t)e * 4);
    const TetConsts kc = tet_consts(I, e, mode == 0);
    for (int j = threadIdx.x; j < U; j += MPS_PATCH) {
        const size_t n = (size_t)I.pnd[base + j];
        const double x0 = I.X[n], x1 = I.X[N + n], x2 = I.X[2 * N + n];
        const double u0 = I.u[n], u1 = I.u[N + n], u2 = I.u[2 * N + n];
        double2* s = reinterpret_cast<double2*>(sxu + 6 * j);
        s[0] = make_double2(x0, x1);
        s[1] = make_double2(x2, u0);
        s[2] = make_double2(u1, u2);
    }
    pl.store(spc, nullptr, ne * 4);  // the list arrives with the nodes; stored now, it holds no registers below
    __syncthreads();
    unsigned long long dmin = INF_BITS;
    double env[3] = {0.0, 0.0, 0.0};
    T f[4][3];
    if (i < ne) {
        const int l[4] = {(int)(l4.x & 0xffffu), (int)(l4.x >> 16), (int)(l4.y & 0xffffu), (int)(l4.y >> 16)};
        // FP64: the stable-dt length first, from the deformed positions (the reference's operations), then the
        // element from a second read of the staged nodes (the memory clobber keeps the reads apart), so that the
        // 24 node values are not held in registers through the element work
        double dt64 = bits2d(INF_BITS);
        if (sizeof(T) == 8 && mode == 0) {
            double xd[4][3];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                double xa[3], ua[3];
                staged_node(sxu, l[a], xa, ua);
#pragma unroll
                for (int d = 0; d < 3; d++) xd[a][d] = xa[d] + ua[d];
            }
            // dtsk is the same in every lane; the elements of a wave whose dt is certainly above it skip the
            // strict length (tet_dt_above, dt_certify)
            if (!(dtsk < __builtin_inf()) || __any(!tet_dt_above(xd, kc.l0, kc.c, dtsk)))
                dt64 = elem_dt(tet_char_length(xd), kc.l0, kc.c);
            asm volatile("" ::: "memory");
        }
        const bool core_dt = sizeof(T) == 4 && mode == 0;
        double x[4][3], uu[4][3];
#pragma unroll
        for (int a = 0; a < 4; a++) staged_node(sxu, l[a], x[a], uu[a]);
        double dte;
        if constexpr (EN) {
            double r[4];
            dte = tet4_core<T, LAW, false, true>(I, e, nullptr, x, uu, kc, f, core_dt, 0.0, r);
            env[0] = r[0];
            env[1] = r[1];
            if (eflags & MPS_E_KIN) {
                const int nd[4] = {I.pnd[base + l[0]], I.pnd[base + l[1]], I.pnd[base + l[2]], I.pnd[base + l[3]]};
                env[2] = tet_kinetic(I, nd, I.rho ? I.rho[e] : I.rho0, r[3]);
            }
        } else {
            dte = tet4_core<T, LAW, false>(I, e, nullptr, x, uu, kc, f, core_dt, 0.0);
        }
        if (sizeof(T) == 8) dte = dt64;
        dmin = d2bits(dte);
    }
    __syncthreads();  // every element has read its nodes: the staging space takes the forces
    if (i < ne) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) sf[(a * 3 + d) * MPS_PATCH + i] = f[a][d];
    }
    __syncthreads();
    patch_sums<T, 4, MPS_PATCH, true>(I, p, sf, spc, I.prow + base + p);
    if (mode == 0) {
        dmin = block_min_u64(dmin, sh);
        if (threadIdx.x == 0) atomicMin(dt_slot(I, (k + 1) & 1, blockIdx.x & (MPS_DT_SLOTS - 1)), dmin);
        if constexpr (EN) elem_energy_publish(I, env, smem);
        if constexpr (SKIP) dt_certify<4, false>(I, k, dtsk, sh);
    }
}

// Persistent batched kernels walk the work list XCD-aware: with gridDim.x a multiple of 8, the
// blocks of XCD x (= blockIdx.x % 8) stride over the contiguous chunk [x*n/8, (x+1)*n/8) of items,
// so neighbouring items of an instance share one L2.  Any other grid falls back to a plain stride.
__device__ __forceinline__ int batch_first(int n) {
    if (gridDim.x & 7) return blockIdx.x;
    const int x = blockIdx.x & 7;
    return (int)(((long long)n * x) >> 3) + (blockIdx.x >> 3);
}
__device__ __forceinline__ int batch_end(int n) {
    if (gridDim.x & 7) return n;
    const int x = blockIdx.x & 7;
    return (int)(((long long)n * (x + 1)) >> 3);
}
__device__ __forceinline__ int batch_stride() { return (gridDim.x & 7) ? gridDim.x : (gridDim.x >> 3); }

// Persistent batched element kernel: a work item is one patch of one instance (begin = its
// first element); BS work-items share the patch's elements.  At least 2 waves per SIMD: the FP64 hex
// instantiations otherwise take more than 256 registers and run one wave per SIMD (10% slower for
// a batch of FP64 cubes, against a small spill).
template <typename T, int NEN, int LAW, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void k_elem_batch(const MpsInst* __restrict__ insts,
                                                    const MpsWork* __restrict__ work, const int* __restrict__ nwork_p) {
    __shared__ T sf[NEN * 3 * MPS_PATCH];
    __shared__ unsigned long long sh[BS / 64 + 1];
    const int nwork = *nwork_p;
    for (int w = batch_first(nwork); w < batch_end(nwork); w += batch_stride()) {
        const MpsWork W = work[w];
        const MpsInst& I = insts[W.inst];
        const long long k = I.ctl->A;
        if (W.first && threadIdx.x == 0) I.ctl->B = k;
        if (I.ctl->slot[k & 1].go) {
            unsigned long long m = patch_forces<T, NEN, LAW, BS>(I, W.begin / MPS_PATCH, true, sf, nullptr, nullptr);
            m = block_min_u64(m, sh);
            if (threadIdx.x == 0) atomicMin(dt_slot(I, (k + 1) & 1, w & (MPS_DT_SLOTS - 1)), m);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Node kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void write_hist_step(const MpsInst& I, long long k, const MpsStepSlot& s) {
    if (I.hist_cap > 0) {
        size_t idx = (size_t)((k - 1) % I.hist_cap);
        I.hist[idx] = s.time;
        I.hist[(size_t)I.hist_cap + idx] = s.dt;
    }
}

// Half-step velocities w of node n as the step's node kernel receives them (nflags): from v and a with the dt of
// the step (MPS_N_IN_VA: the first step of a launch), from a's array (MPS_N_IN_WA: after a step that kept v
// readable) or from v's array.
__device__ __forceinline__ void load_half_velocity(const MpsInst& I, int n, unsigned nflags, double dt, double* w) {
    const size_t N = (size_t)I.N;
    if (nflags & MPS_N_IN_VA) {
#pragma unroll
        for (int i = 0; i < 3; i++) w[i] = half_velocity(dt, I.v[i * N + n], I.a[i * N + n]);
    } else {
        const double* src = (nflags & MPS_N_IN_WA) ? I.a : I.v;
#pragma unroll
        for (int i = 0; i < 3; i++) w[i] = src[i * N + n];
    }
}

// Node update of step k (node kernel).  nflags: where step k-1 left w (load_half_velocity); MPS_N_OUT_V, keep v
// of step k readable for the next element phase (kinetic energy), w then going to a's array.  Without a
// prediction (fuse = 0, or no step k+1) v and a are kept, so that the state between launches is u, v, a.  EN: energy bookkeeping of the node into red[MPS_NODE_PARTS]
// (node_energy_forces, node_energy_motion).  Returns false when the step is not running (the same for every
// work-item of the launch).
template <bool EN = false>
__device__ __forceinline__ bool node_item(const MpsInst& I, int n, bool first, int fuse, unsigned nflags,
                                          double* red = nullptr) {
    MpsCtl* ctl = I.ctl;
    const long long k = ctl->B;
    const MpsStepSlot s = ctl->slot[k & 1];
    if (!s.go) return false;
    MpsStepSlot nx;
    nx.go = 0;
    int st = MPSD_RUNNING;
    const unsigned long long mb = fuse ? dt_read_min(I, (k + 1) & 1) : 0ull;
    if (fuse) st = begin_scalars(I, k + 1, s, mb, &nx);
    if (first) {
        ctl->A = k + 1;
        dt_reset(I, k & 1);
        if (fuse) {
            ctl->dt_skip = dt_skip_next(I, k + 1, mb);
            ctl->slot[(k + 1) & 1] = nx;
            if (st != MPSD_RUNNING) ctl->status = st;
            if (nx.go) write_hist_step(I, k + 1, nx);
        }
    }
    if (n < 0) return true;
    const size_t N = (size_t)I.N;
    double u[3], w[3], v[3], a[3];
#pragma unroll
    for (int i = 0; i < 3; i++) u[i] = I.u[i * N + n];
    load_half_velocity(I, n, nflags, s.dt, w);
    if (EN) {
        // the forces first (also of a fully constrained node, for the work of the supports), then the step's
        // increments from w: nothing extra is held through the partial sums
        const uint32_t info = I.ninfo[n];
        double rhs[3];
        {
            double fi[3], fe[3], incr[3];
            node_partial_sum(I, n, fi);
            node_fext(I, n, info, fe);
#pragma unroll
            for (int i = 0; i < 3; i++) incr[i] = step_increment(I, info, i, k, s.dt, w[i]);
            node_energy_forces(I, n, info, k, u, incr, fi, fe, red);
            node_rhs(fi, fe, rhs);
        }
        const double m = I.mass[n];
        double wkin = -0.0;
        node_update(I, info, k, s.dt, m, rhs, w, v, a, &wkin);
        node_energy_motion(m, k, wkin, v, red);
    } else {
        solve_node(I, n, k, s.dt, w, v, a);
    }
    if (n == I.dof1_node) {
        double x = u[I.dof1_dir];
        if (!isfinite(x)) {
            ctl->nonfinite = 1;
            atomicMin((unsigned long long*)&ctl->nan_step, (unsigned long long)k);
        }
    }
    const bool predict = fuse && nx.go;
    if (predict) {
        if (I.uprev) {  // the F of step k becomes the previous state of step k+1 (nextStep)
#pragma unroll
            for (int i = 0; i < 3; i++) I.uprev[i * N + n] = u[i];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) w[i] = half_velocity(nx.dt, v[i], a[i]);
        predict_node(I, n, k + 1, nx.dt, u, w);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) I.u[i * N + n] = u[i];
    if (!predict) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            I.v[i * N + n] = v[i];
            I.a[i * N + n] = a[i];
        }
    } else if (nflags & MPS_N_OUT_V) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            I.v[i * N + n] = v[i];
            I.a[i * N + n] = w[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) I.v[i * N + n] = w[i];
    }
    return true;
}

// Node sums of step j (parity j&1 of npart, nb blocks) into the running works and the history, once per step
// (ctl->wred).  Called by every thread of one block.
__device__ __forceinline__ void node_sums_history(const MpsInst& I, long long j, const double* npart, int nb,
                                                  double* sh) {
    __shared__ int todo;
    MpsCtl* ctl = I.ctl;
    if (threadIdx.x == 0) todo = (j >= 1 && ctl->wred < j) ? 1 : 0;
    __syncthreads();
    if (!todo) return;
    double t[MPS_NODE_PARTS];
    reduce_parts<MPS_NODE_PARTS>(npart + (size_t)(j & 1) * MPS_NODE_PARTS * nb, nb, t, sh);
    if (threadIdx.x == 0) {
        ctl->w_ext += t[0];
        ctl->w_react += t[1];
        *hist_at(I, MPS_H_WEXT, j) = ctl->w_ext;
        *hist_at(I, MPS_H_WREACT, j) = ctl->w_react;
        if (j % I.energy_every == 0) {
            *hist_at(I, MPS_H_EKL, j) = t[2];
            *hist_at(I, MPS_H_EEXT, j) = t[3];
        }
        ctl->wred = j;
    }
}

// fuse = 1: also begin step k+1.  EN (energies on): each block's work and energy sums go to epart (parity k&1),
// and block 0 adds up what earlier kernels of this launch left there: the element phase of step k (nflags
// MPS_E_DEF: DEFO_ENERGY and VOLUME of step k; MPS_E_KIN: KIN_ENERGY of step k-1) and, with MPS_N_PREV, the
// node sums of step k-1.
template <bool EN>
__global__ __launch_bounds__(256) void k_node(MpsInst I, int fuse, unsigned nflags) {
    const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool first = (n == 0);
    if (!EN) {
        node_item(I, n < I.N ? n : -1, first, fuse, nflags);
        return;
    }
    __shared__ double sh[MPS_NODE_PARTS * 256];
    // sums start at -0.0, the identity of FP64 addition (0.0 + x is not x for x = -0.0), so that the compiler
    // can drop the first additions and hold no zeros through the node work
    double red[MPS_NODE_PARTS] = {-0.0, -0.0, -0.0, -0.0};
    const long long k = I.ctl->B;  // read before the first work-item advances the step counters
    if (!node_item<true>(I, n < I.N ? n : -1, first, fuse, nflags, red)) return;
    const int nb = (int)gridDim.x;
    double* npart = I.epart + 3 * (size_t)I.n_patch;
    publish_block_sum<MPS_NODE_PARTS>(red, npart + (size_t)(k & 1) * MPS_NODE_PARTS * nb, sh);
    if (blockIdx.x != 0 || I.hist_cap <= 0) return;
    if (nflags & (MPS_E_DEF | MPS_E_KIN)) {
        double t[3];
        reduce_parts<3>(I.epart, I.n_patch, t, sh);
        if (threadIdx.x == 0) {
            if (nflags & MPS_E_DEF) {
                *hist_at(I, MPS_H_EDEF, k) = t[0];
                *hist_at(I, MPS_H_VOL, k) = t[1];
            }
            if ((nflags & MPS_E_KIN) && k >= 2) *hist_at(I, MPS_H_EKIN, k - 1) = t[2];
        }
    }
    if (nflags & MPS_N_PREV) node_sums_history(I, k - 1, npart, nb, sh);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_node_batch(const MpsInst* __restrict__ insts, const MpsWork* __restrict__ work,
                                                    const int* __restrict__ nwork_p, int fuse, unsigned nflags) {
    const int nwork = *nwork_p;
    for (int w = batch_first(nwork); w < batch_end(nwork); w += batch_stride()) {
        const MpsWork W = work[w];
        const MpsInst& I = insts[W.inst];
        const int n = W.begin + threadIdx.x;
        const bool first = W.first && threadIdx.x == 0;
        node_item(I, n < W.end ? n : -1, first, fuse, nflags);
    }
}

// Per-step DOF series (mps_set_dof_series), after the element kernel (and the loads) of step k = ctl->B and
// before its node kernel: u holds u_k there, part the step's partial forces, lf its external forces.  The
// velocity and acceleration are those the node kernel is about to give the DOF (solve_node on a copy, the same
// arithmetic, from the state as the node kernel receives it: nflags as for it), the internal force
// the node's partials in CSR order as the node kernel sums them.
__global__ __launch_bounds__(256) void k_series(MpsInst I, unsigned nflags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I.n_series) return;
    const long long k = I.ctl->B;
    const MpsStepSlot s = I.ctl->slot[k & 1];
    if (!s.go) return;
    const size_t N = (size_t)I.N;
    const int dof = I.ser_dof[i], kind = I.ser_kind[i];
    const int n = dof % I.N, dir = dof / I.N;
    double val;
    if (kind == MPS_SERIES_U) {
        val = I.u[dof];
    } else if (kind == MPS_SERIES_FINT) {
        val = 0.0;
        for (int q = I.nrow[n]; q < I.nrow[n + 1]; q++) {
            double c[3];
            load_partial(I, I.ncol[q], c);
            val += c[dir];
        }
    } else if (kind == MPS_SERIES_FEXT) {
        double fe[3];
        node_fext(I, n, I.ninfo[n], fe);
        val = fe[dir];
    } else {
        double w[3], v[3], a[3];
        load_half_velocity(I, n, nflags, s.dt, w);
        solve_node(I, n, k, s.dt, w, v, a);
        val = kind == MPS_SERIES_V ? v[dir] : a[dir];
    }
    I.ser_ring[(size_t)((k - 1) % I.hist_cap) * I.n_series + i] = val;
}

// ---------------------------------------------------------------------------
// Begin of step k = A (first step of a launch): the prediction from v and a, which stay in place for the
// step's node kernel (MPS_N_IN_VA)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void begin_item(const MpsInst& I, int n, bool first) {
    MpsCtl* ctl = I.ctl;
    const long long k = ctl->A;
    const MpsStepSlot prev = ctl->slot[(k - 1) & 1];
    MpsStepSlot nx;
    const unsigned long long mb = dt_read_min(I, k & 1);
    const int st = begin_scalars(I, k, prev, mb, &nx);
    if (first) {
        ctl->dt_skip = dt_skip_next(I, k, mb);
        ctl->slot[k & 1] = nx;
        if (st != MPSD_RUNNING) ctl->status = st;
        if (nx.go) write_hist_step(I, k, nx);
    }
    if (n < 0 || !nx.go) return;
    const size_t N = (size_t)I.N;
    double u[3], w[3];
#pragma unroll
    for (int i = 0; i < 3; i++) u[i] = I.u[i * N + n];
    load_half_velocity(I, n, MPS_N_IN_VA, nx.dt, w);
    if (I.uprev) {
#pragma unroll
        for (int i = 0; i < 3; i++) I.uprev[i * N + n] = u[i];
    }
    predict_node(I, n, k, nx.dt, u, w);
#pragma unroll
    for (int i = 0; i < 3; i++) I.u[i * N + n] = u[i];
}

__global__ __launch_bounds__(256) void k_begin(MpsInst I) {
    const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    begin_item(I, n < I.N ? n : -1, n == 0);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_begin_batch(const MpsInst* __restrict__ insts, const MpsWork* __restrict__ work,
                                                     const int* __restrict__ nwork_p) {
    const int nwork = *nwork_p;
    for (int w = batch_first(nwork); w < batch_end(nwork); w += batch_stride()) {
        const MpsWork W = work[w];
        const int n = W.begin + threadIdx.x;
        begin_item(insts[W.inst], n < W.end ? n : -1, W.first && threadIdx.x == 0);
    }
}

// ---------------------------------------------------------------------------
// Stable dt over all elements for a given u (NULL: undeformed), written as a
// 64-bit pattern min into *out (solvers.cpp:276-320).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dt(MpsInst I, const double* u, unsigned long long* out) {
    __shared__ unsigned long long sh[4];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double dte = bits2d(INF_BITS);
    if (e < I.Ne) {
        double pa, pb, c;
        elem_mat(I, e, pa, pb, c);
        const size_t N = (size_t)I.N;
        if (u == nullptr) {
            dte = elem_dt(I.l0[e], I.l0[e], c);
        } else if (I.nen == 8) {
            double x[8][3];
            for (int a = 0; a < 8; a++) {
                int nd = I.conn[(size_t)e * 8 + a];
                for (int d = 0; d < 3; d++) x[a][d] = I.X[d * N + nd] + u[d * N + nd];
            }
            dte = elem_dt(hex_char_length(x), I.l0[e], c);
        } else {
            double x[4][3];
            for (int a = 0; a < 4; a++) {
                int nd = I.conn[(size_t)e * 4 + a];
                for (int d = 0; d < 3; d++) x[a][d] = I.X[d * N + nd] + u[d * N + nd];
            }
            dte = elem_dt(tet_char_length(x), I.l0[e], c);
        }
    }
    unsigned long long m = block_min_u64(d2bits(dte), sh);
    if (threadIdx.x == 0) atomicMin(out, m);
}

// ---------------------------------------------------------------------------
// Set-up: reference-order Gauss point geometry (classLinear*::defineGPs), initial
// characteristic length, lumped-mass slot values (constitutiveAlgorithms.cpp:191-205),
// min l0/c.  No FMA contraction: the lumped mass is bitwise the reference's
// per-element sums.
// ---------------------------------------------------------------------------

// Shape values, natural derivatives and weight of Gauss point g of the set-up quadrature: 2x2x2 for
// hex (GPsDistribution.cpp:392-432), one point for tet (:339-343).
template <int NEN>
__device__ __forceinline__ void gp_shape(int g, double* phi, double* nd, double* weight) {
#pragma clang fp contract(off)
    if (NEN == 8) {
        const double gx = sxi(g) > 0 ? MPS_GP : -MPS_GP;
        const double gy = seta(g) > 0 ? MPS_GP : -MPS_GP;
        const double gz = szeta(g) > 0 ? MPS_GP : -MPS_GP;
        hex_shape_ref(gx, gy, gz, phi, nd);
        *weight = 1 * 1.0;
    } else {
        const double q = 0.250000000000, r = 0.25;
        phi[0] = -1.0 * q - 1.0 * r - 1.0 * r + 1.0;
        phi[1] = 1.0 * q;
        phi[2] = 1.0 * r;
        phi[3] = 1.0 * r;
        for (int k = 0; k < 12; k++) nd[k] = 0.0;
        nd[0] = -1.0 - 0.0 - 0.0 + 0.0;
        nd[3] = 1.0;
        nd[1] = -0.0 - 1.0 - 0.0 + 0.0;
        nd[7] = 1.0;
        nd[2] = -0.0 - 0.0 - 1.0 + 0.0;
        nd[11] = 1.0;
        *weight = MPS_TET_W;
    }
}

// det of the Jacobian xyzCoor * natDev, k innermost (maths.cpp:1679-1693).
template <int NEN>
__device__ __forceinline__ double gp_det(const double (*x)[3], const double* nd) {
#pragma clang fp contract(off)
    double Jac[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < NEN; k++) s += x[k][i] * nd[k * 3 + j];
            Jac[i * 3 + j] = s;
        }
    return det3_strict(Jac);
}

template <int NEN>
__device__ __forceinline__ void load_ref_coords(const MpsInst& I, int e, double (*x)[3]) {
    const size_t N = (size_t)I.N;
    for (int a = 0; a < NEN; a++) {
        const int nd = I.conn[(size_t)e * NEN + a];
        for (int d = 0; d < 3; d++) x[a][d] = I.X[d * N + nd];
    }
}

template <int NEN>
__device__ __forceinline__ void setup_element(const MpsInst& I, int e, double rho, double* ms, double* l0) {
#pragma clang fp contract(off)
    double x[NEN][3];
    load_ref_coords<NEN>(I, e, x);
    *l0 = (NEN == 8) ? hex_char_length((const double(*)[3])x) : tet_char_length((const double(*)[3])x);
    for (int a = 0; a < NEN; a++) ms[a] = 0.0;
    const int ngp = (NEN == 8) ? 8 : 1;
    for (int g = 0; g < ngp; g++) {
        double phi[8], nd[24];
        double weight;
        gp_shape<NEN>(g, phi, nd, &weight);
        double wJ = weight * fabs(gp_det<NEN>((const double(*)[3])x, nd));
        for (int a = 0; a < NEN; a++) {
            double s = 0.0;
            for (int b = 0; b < NEN; b++) s += wJ * rho * phi[a] * phi[b];
            ms[a] += s;
        }
    }
}

__global__ __launch_bounds__(256) void k_setup_elem(MpsInst I, double* __restrict__ mslot, double* __restrict__ l0out,
                                                    unsigned long long* dt0) {
    __shared__ unsigned long long sh[4];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double dte = bits2d(INF_BITS);
    if (e < I.Ne) {
        double pa, pb, c;
        elem_mat(I, e, pa, pb, c);
        const double rho = I.rho ? I.rho[e] : I.rho0;
        double l0;
        if (I.nen == 8) setup_element<8>(I, e, rho, mslot + (size_t)e * 8, &l0);
        else setup_element<4>(I, e, rho, mslot + (size_t)e * 4, &l0);
        l0out[e] = l0;
        dte = elem_dt(l0, l0, c);
    }
    unsigned long long m = block_min_u64(d2bits(dte), sh);
    if (threadIdx.x == 0) atomicMin(dt0, m);
}

// Initial control block of a solve from the set-up reduction (device side, so that admitting
// a solve into a running batch needs no host round trip): slot[0] = {t0, s*min(l0/c), go}.
__global__ void k_fill_u64(unsigned long long* p, long long n, unsigned long long v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_init_ctl(MpsInst I, const unsigned long long* dt0) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    MpsCtl* c = I.ctl;
    c->slot[0].time = I.t0;
    c->slot[0].dt = bits2d(*dt0) * I.scale;  // solverExplicit.cpp:165-166
    c->slot[0].go = 1;
    c->slot[1].time = I.t0;
    c->slot[1].dt = 0.0;
    c->slot[1].go = 0;
    c->A = 1;
    c->B = 0;
    c->status = MPSD_RUNNING;
    c->nonfinite = 0;
    c->nan_step = 0x7FFFFFFFFFFFFFFFll;
    c->dt0min = *dt0;
    c->w_ext = 0.0;
    c->w_react = 0.0;
    c->wred = 0;
    c->dt_skip = __builtin_inf();
    c->skip_hold = 0;
}

// Node-centric gather of per-(element,slot) scalars: out[n] = sum over CSR(n), ascending.
__global__ __launch_bounds__(256) void k_gather_scalar(MpsInst I, const double* __restrict__ slotv, double* out) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= I.N) return;
    double acc = 0.0;
    for (int q = I.rowptr[n]; q < I.rowptr[n + 1]; q++) acc += slotv[I.col[q]];
    out[n] = acc;
}

// Gravity load (BCsGravity, boundaryConditions.cpp:1146-1172), constant over the run: per Gauss point the
// factor -w * rho * |J| * g (tslot[e*ngp + g]), then per node the sum of factor * phi_a over its
// elements in ascending order and their Gauss points in order, as the reference adds them.
template <int NEN>
__global__ __launch_bounds__(256) void k_grav_elem(MpsInst I, double grav, double* __restrict__ tslot) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= I.Ne) return;
    const double rho = I.rho ? I.rho[e] : I.rho0;
    double x[NEN][3];
    load_ref_coords<NEN>(I, e, x);
    const int ngp = (NEN == 8) ? 8 : 1;
    for (int g = 0; g < ngp; g++) {
        double phi[8], nd[24], w;
        gp_shape<NEN>(g, phi, nd, &w);
        const double J = fabs(gp_det<NEN>((const double(*)[3])x, nd));
        tslot[(size_t)e * ngp + g] = -w * rho * J * grav;
    }
}

template <int NEN>
__global__ __launch_bounds__(256) void k_grav_node(MpsInst I, const double* __restrict__ tslot, double* fgrav) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= I.N) return;
    const int ngp = (NEN == 8) ? 8 : 1;
    double acc = 0.0;
    for (int q = I.rowptr[n]; q < I.rowptr[n + 1]; q++) {
        const int slot = I.col[q];
        const int e = slot / NEN, a = slot - e * NEN;
        for (int g = 0; g < ngp; g++) {
            double phi[8], nd[24], w;
            gp_shape<NEN>(g, phi, nd, &w);
            acc += tslot[(size_t)e * ngp + g] * phi[a];
        }
    }
    fgrav[n] = acc;
}

// Surface forces of every pressure Gauss point: mode 0 at the time of the step in flight (ctl->B, after its
// element kernel), mode 1 at time t (standalone evaluation).
__global__ __launch_bounds__(256) void k_press(MpsInst I, int mode, double t) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        const MpsStepSlot s = I.ctl->slot[I.ctl->B & 1];
        if (!s.go) return;
        t = s.time;
    }
    if (q < I.n_sgp) surface_point(I, q, t);
}

// External force of every flagged node: mode 0 at the time of the step in flight (after its surface
// kernel), mode 1 at time t.
__global__ __launch_bounds__(256) void k_loads(MpsInst I, int mode, double t) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        const MpsStepSlot s = I.ctl->slot[I.ctl->B & 1];
        if (!s.go) return;
        t = s.time;
    }
    if (j >= I.n_lnode) return;
    double fe[3];
    loaded_node_fext(I, j, t, fe);
    for (int d = 0; d < 3; d++) I.lf[3 * (size_t)j + d] = fe[d];
}

// External force of every DOF (NeumannBCManagement, boundaryConditions.cpp:75-140) from the values of k_loads.
__global__ __launch_bounds__(256) void k_fext(MpsInst I, double* fext) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= I.N) return;
    double fe[3];
    node_fext(I, n, I.ninfo[n], fe);
    const size_t N = (size_t)I.N;
    fext[n] = fe[0];
    fext[N + n] = fe[1];
    fext[2 * N + n] = fe[2];
}

// Internal force for all DOFs (solverExplicit.cpp:363-379), from the patch partials.
__global__ __launch_bounds__(256) void k_gather_fint(MpsInst I, double* fint) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= I.N) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int q = I.nrow[n]; q < I.nrow[n + 1]; q++) {
        double c[3];
        load_partial(I, I.ncol[q], c);
        f0 += c[0];
        f1 += c[1];
        f2 += c[2];
    }
    const size_t N = (size_t)I.N;
    fint[n] = f0;
    fint[N + n] = f1;
    fint[2 * N + n] = f2;
}

// ---------------------------------------------------------------------------
// Energies, FP64 in every form (DEFO_ENERGY, KIN_ENERGY, VOLUME with the Sum operation, dataExtraction.cpp:
// 569-776), with the arithmetic of the fused element kernels: per hex8 GP J, cof(J), wJ = |det J| and
// H = grad_xi(u) cof(J)^T / det J from the butterfly coefficients, W and det F from pk1, the kinetic energy from
// the butterfly coefficients of v (hex_kinetic); per tet the single GP of tet4_core.  MPS_E_DEF: DEFO_ENERGY and
// VOLUME of the current u; MPS_E_KIN / MPS_E_KIN_NOW: KIN_ENERGY of the current v.
// ---------------------------------------------------------------------------
template <int LAW>
__device__ __forceinline__ void elem_energy(const MpsInst& I, int e, unsigned flags, double* out) {
    const size_t N = (size_t)I.N;
    double pa, pb, c;
    elem_mat(I, e, pa, pb, c);
    const double rho = I.rho ? I.rho[e] : I.rho0;
    double ed = 0.0, ek = 0.0, vol = 0.0;
    if (I.nen == 8) {
        int nd[8];
#pragma unroll
        for (int a = 0; a < 8; a++) nd[a] = I.conn[(size_t)e * 8 + a];
        double Ax[7][3], Au[7][3];
        {
            double x[8][3], cx[7][3];
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int d = 0; d < 3; d++) x[a][d] = I.X[d * N + nd[a]];
            hex_butterfly(x, cx);
            hex_scale<double>(cx, Ax);
            if (flags & MPS_E_DEF) {
#pragma unroll
                for (int a = 0; a < 8; a++)
#pragma unroll
                    for (int d = 0; d < 3; d++) x[a][d] = I.u[d * N + nd[a]];
                hex_butterfly(x, cx);
                hex_scale<double>(cx, Au);
            }
        }
        double wJ[8];
#pragma unroll 1
        for (int gp = 0; gp < 8; gp++) {
            const GpSigns<double> s = gp_signs<double>(gp);
            double Jm[9], cJ[9];
            hex_jac<double>(Ax, s, Jm);
            cof3(Jm, cJ);
            const double dJ = Jm[0] * cJ[0] + Jm[1] * cJ[1] + Jm[2] * cJ[2];
            const double w = fabs(dJ);
            wJ[gp] = w;
            if (flags & MPS_E_DEF) {
                double Du[9], H[9], P[9], Wg, Jg;
                hex_jac<double>(Au, s, Du);
                const double rJ = rcp(dJ);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        H[i * 3 + j] =
                            (Du[i * 3] * cJ[j * 3] + Du[i * 3 + 1] * cJ[j * 3 + 1] + Du[i * 3 + 2] * cJ[j * 3 + 2]) * rJ;
                pk1<LAW>(H, pa, pb, P, &Wg, &Jg);
                ed += Wg * w;
                vol += w * Jg;
            }
        }
        if (flags & (MPS_E_KIN | MPS_E_KIN_NOW)) ek = hex_kinetic(I, nd, rho, wJ);
    } else {
        const int4 q = *reinterpret_cast<const int4*>(I.conn + (size_t)e * 4);
        const int nd[4] = {q.x, q.y, q.z, q.w};
        double x[4][3], uu[4][3], f[4][3], r[4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                x[a][d] = I.X[d * N + nd[a]];
                uu[a][d] = I.u[d * N + nd[a]];
            }
        tet4_core<double, LAW, false, true>(I, e, nd, x, uu, tet_consts(I, e, false), f, false, 0.0, r);
        ed = r[0];
        vol = r[1];
        if (flags & (MPS_E_KIN | MPS_E_KIN_NOW)) ek = tet_kinetic(I, nd, rho, r[3]);
    }
    out[0] = ed;
    out[1] = vol;
    out[2] = ek;
}

// ---------------------------------------------------------------------------
// Element fields of the VTK output (output.cpp:567-820): per element the GP averages of the
// state quantities updateConstitutive leaves (Cauchy, Green-Lagrange strain -- the small strain for
// the linear law --, von Mises stress, volumetric and equivalent strain; maths.cpp:1241-1299).
// Quantities a law does not set are 0 (classStates.cpp:15-44).  FP64 in every precision form.
// ---------------------------------------------------------------------------
__device__ void gp_outputs(int law, const double* F, double pa, double pb, double* q) {
    double H[9], P[9];
    double* cau = q;
    double* E = q + 9;
#pragma unroll
    for (int i = 0; i < 21; i++) q[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) H[i] = F[i] - ((i % 4) == 0 ? 1.0 : 0.0);
    if (law == 1 || law == 2) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                E[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i] + H[i] * H[j] + H[3 + i] * H[3 + j] + H[6 + i] * H[6 + j]);
    } else if (law == 3) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) E[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    }
    const double J = det3(F);
    if (law == 1) {  // sigma = mu J^{-5/3} B + (-(mu/3) tr B J^{-5/3} + K (J-1)) I
        double B[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) B[i * 3 + j] = F[i * 3] * F[j * 3] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2];
        const double f1 = pow(J, 5.0 / 3.0);
        const double f3 = -(pa / 3.0) * (B[0] + B[4] + B[8]) / f1 + pb * (J - 1.0);
        for (int i = 0; i < 9; i++) cau[i] = (pa / f1) * B[i] + ((i % 4) == 0 ? f3 : 0.0);
    } else {
        if (law == 0) pk1<0>(H, pa, pb, P);
        else if (law == 2) pk1<2>(H, pa, pb, P);
        else pk1<3>(H, pa, pb, P);
        const double rJ = 1.0 / J;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                cau[i * 3 + j] = (P[i * 3] * F[j * 3] + P[i * 3 + 1] * F[j * 3 + 1] + P[i * 3 + 2] * F[j * 3 + 2]) * rJ;
    }
    q[18] = sqrt(0.5 * ((cau[0] - cau[4]) * (cau[0] - cau[4]) + (cau[4] - cau[8]) * (cau[4] - cau[8]) +
                        (cau[8] - cau[0]) * (cau[8] - cau[0]) + 6.0 * (cau[1] * cau[1] + cau[2] * cau[2] + cau[5] * cau[5])));
    if (law == 1 || law == 2) {
        q[19] = E[0] + E[4] + E[8];
        q[20] = 0.6666 * sqrt(1.5 * (E[0] * E[0] + E[4] * E[4] + E[8] * E[8]) + 3.0 * (E[1] * E[1] + E[2] * E[2] + E[5] * E[5]));
    }
}

__global__ __launch_bounds__(256) void k_elem_outputs(MpsInst I, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= I.Ne) return;
    const size_t N = (size_t)I.N;
    double pa, pb, c;
    elem_mat(I, e, pa, pb, c);
    double acc[21], q[21];
    for (int i = 0; i < 21; i++) acc[i] = 0.0;
    int ngp;
    if (I.nen == 8) {
        double x[8][3], uu[8][3];
        for (int a = 0; a < 8; a++) {
            const int nd = I.conn[(size_t)e * 8 + a];
            for (int d = 0; d < 3; d++) {
                x[a][d] = I.X[d * N + nd];
                uu[a][d] = I.u[d * N + nd];
            }
        }
        double cx[7][3], cu[7][3], Ax[7][3], Au[7][3];
        hex_butterfly(x, cx);
        hex_butterfly(uu, cu);
        hex_scale<double>(cx, Ax);
        hex_scale<double>(cu, Au);
        for (int gp = 0; gp < 8; gp++) {
            const GpSigns<double> s = gp_signs<double>(gp);
            double Jm[9], Du[9], iJ[9], F[9];
            hex_jac<double>(Ax, s, Jm);
            hex_jac<double>(Au, s, Du);
            inv3(Jm, iJ, det3(Jm));
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    F[i * 3 + j] = (i == j ? 1.0 : 0.0) + Du[i * 3] * iJ[j] + Du[i * 3 + 1] * iJ[3 + j] + Du[i * 3 + 2] * iJ[6 + j];
            gp_outputs(I.law, F, pa, pb, q);
            for (int i = 0; i < 21; i++) acc[i] += q[i];
        }
        ngp = 8;
    } else {
        double x[4][3], uu[4][3];
        for (int a = 0; a < 4; a++) {
            const int nd = I.conn[(size_t)e * 4 + a];
            for (int d = 0; d < 3; d++) {
                x[a][d] = I.X[d * N + nd];
                uu[a][d] = I.u[d * N + nd];
            }
        }
        double Jm[9], Du[9], iJ[9], F[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Jm[i * 3 + j] = x[j + 1][i] - x[0][i];
                Du[i * 3 + j] = uu[j + 1][i] - uu[0][i];
            }
        inv3(Jm, iJ, det3(Jm));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                F[i * 3 + j] = (i == j ? 1.0 : 0.0) + Du[i * 3] * iJ[j] + Du[i * 3 + 1] * iJ[3 + j] + Du[i * 3 + 2] * iJ[6 + j];
        gp_outputs(I.law, F, pa, pb, acc);
        ngp = 1;
    }
    const size_t eo = I.eorig ? (size_t)I.eorig[e] : (size_t)e;  // rows in input element order
    for (int i = 0; i < 21; i++) out[eo * 21 + i] = acc[i] / (double)ngp;
}

// Standalone energy pass: between the element and node kernels of step k = ctl->B (flags MPS_E_DEF: DEFO_ENERGY
// and VOLUME of step k; MPS_E_KIN: KIN_ENERGY of step k-1, whose velocity v still holds), or after the last node
// kernel of a launch (MPS_E_KIN_NOW: KIN_ENERGY of the last step that ran, k = ctl->A - 1, when it is an energy
// step -- in the time-window mode the solve may end before the launch's last step).
template <int LAW>
__global__ __launch_bounds__(MPS_PATCH) void k_energy(MpsInst I, unsigned flags) {
    __shared__ double sh[3 * MPS_PATCH];
    const long long k = (flags & MPS_E_KIN_NOW) ? I.ctl->A - 1 : I.ctl->B;
    if (k < 1 || !I.ctl->slot[k & 1].go) return;
    if ((flags & MPS_E_KIN_NOW) && k % I.energy_every != 0) return;
    const int e = blockIdx.x * MPS_PATCH + threadIdx.x;  // n_patch blocks: the element phase's block sums
    double vals[3] = {0.0, 0.0, 0.0};
    if (e < I.Ne) elem_energy<LAW>(I, e, flags, vals);
    publish_block_sum<3>(vals, I.epart, sh);
}

// End of a launch (one block), after the MPS_E_KIN_NOW pass: the node sums of the last step that ran,
// k = ctl->A - 1, and its KIN_ENERGY when it is an energy step.
__global__ __launch_bounds__(256) void k_energy_tail(MpsInst I) {
    __shared__ double sh[MPS_NODE_PARTS * 256];
    const long long k = I.ctl->A - 1;
    if (k < 1 || I.hist_cap <= 0) return;
    node_sums_history(I, k, I.epart + 3 * (size_t)I.n_patch, (I.N + 255) / 256, sh);
    if (k % I.energy_every == 0 && I.ctl->slot[k & 1].go) {  // the conditions of the pass before
        double t[3];
        reduce_parts<3>(I.epart, I.n_patch, t, sh);
        if (threadIdx.x == 0) *hist_at(I, MPS_H_EKIN, k) = t[2];
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
namespace mpsk {

static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

bool elem_fuses_energy(const MpsInst& I) { return I.precision == 0 && I.uprev == nullptr; }

// Dynamic LDS of the staged tet4 kernel: node staging / force slots, then the contribution list.
static size_t tet_lds(const MpsInst& I) { return tet_shared_bytes(I.pu_max, I.precision == 1 ? 4 : 8) + 8 * MPS_PATCH; }

hipError_t launch_elem(const MpsInst& I, int mode, hipStream_t st, unsigned eflags) {
    const unsigned nb = (unsigned)I.n_patch;
    if (nb == 0) return hipSuccess;
    const bool visc = mode == 0 && I.uprev != nullptr;
    const bool en = mode == 0 && eflags != 0u && elem_fuses_energy(I);
#define MPS_ELEM_LAW_O(T, NEN, V, E, O)                                                          \
    switch (I.law) {                                                                             \
        case 0: k_elem<T, NEN, 0, V, E, O><<<nb, MPS_PATCH, 0, st>>>(I, mode, eflags); break;    \
        case 1: k_elem<T, NEN, 1, V, E, O><<<nb, MPS_PATCH, 0, st>>>(I, mode, eflags); break;    \
        case 2: k_elem<T, NEN, 2, V, E, O><<<nb, MPS_PATCH, 0, st>>>(I, mode, eflags); break;    \
        default: k_elem<T, NEN, 3, V, E, O><<<nb, MPS_PATCH, 0, st>>>(I, mode, eflags); break;   \
    }
#define MPS_ELEM_LAW(T, NEN, V, E) MPS_ELEM_LAW_O(T, NEN, V, E, false)
    // the inviscid hex8 kernels gather their nodes by 32-bit offsets when the mesh allows it
    const bool o32 = I.off32 != 0;
    if (I.nen == 8) {
        if (I.precision == 1) {
            if (visc) {
                MPS_ELEM_LAW(float, 8, true, false)
            } else if (o32) {
                MPS_ELEM_LAW_O(float, 8, false, false, true)
            } else {
                MPS_ELEM_LAW(float, 8, false, false)
            }
        } else if (visc) {
            MPS_ELEM_LAW(double, 8, true, false)
        } else if (en) {
            if (o32) { MPS_ELEM_LAW_O(double, 8, false, true, true) } else { MPS_ELEM_LAW(double, 8, false, true) }
        } else {
            if (o32) { MPS_ELEM_LAW_O(double, 8, false, false, true) } else { MPS_ELEM_LAW(double, 8, false, false) }
        }
    } else if (!visc && I.pnd != nullptr) {
        const size_t lds = tet_lds(I);
#define MPS_TET_STAGED(T, E)                                                                             \
    switch (I.law) {                                                                                     \
        case 0: k_elem_tet_staged<T, 0, E><<<nb, MPS_PATCH, lds, st>>>(I, mode, eflags); break;          \
        case 1: k_elem_tet_staged<T, 1, E><<<nb, MPS_PATCH, lds, st>>>(I, mode, eflags); break;          \
        case 2: k_elem_tet_staged<T, 2, E><<<nb, MPS_PATCH, lds, st>>>(I, mode, eflags); break;          \
        default: k_elem_tet_staged<T, 3, E><<<nb, MPS_PATCH, lds, st>>>(I, mode, eflags); break;         \
    }
        if (I.precision == 1) {
            MPS_TET_STAGED(float, false)
        } else if (en) {
            MPS_TET_STAGED(double, true)
        } else {
            MPS_TET_STAGED(double, false)
        }
#undef MPS_TET_STAGED
    } else {
        if (I.precision == 1) {
            if (visc) { MPS_ELEM_LAW(float, 4, true, false) } else { MPS_ELEM_LAW(float, 4, false, false) }
        } else if (visc) {
            MPS_ELEM_LAW(double, 4, true, false)
        } else if (en) {
            MPS_ELEM_LAW(double, 4, false, true)
        } else {
            MPS_ELEM_LAW(double, 4, false, false)
        }
    }
#undef MPS_ELEM_LAW
#undef MPS_ELEM_LAW_O
    return hipGetLastError();
}

hipError_t launch_node(const MpsInst& I, int fuse, hipStream_t st, unsigned nflags) {
    if (I.energy_every > 0) k_node<true><<<blocks_for(I.N, 256), 256, 0, st>>>(I, fuse, nflags);
    else k_node<false><<<blocks_for(I.N, 256), 256, 0, st>>>(I, fuse, nflags & (MPS_N_IN_VA | MPS_N_IN_WA | MPS_N_OUT_V));
    return hipGetLastError();
}

hipError_t launch_begin(const MpsInst& I, hipStream_t st) {
    k_begin<<<blocks_for(I.N, 256), 256, 0, st>>>(I);
    return hipGetLastError();
}

hipError_t launch_dt(const MpsInst& I, const double* u, unsigned long long* out, hipStream_t st) {
    k_dt<<<blocks_for(I.Ne, 256), 256, 0, st>>>(I, u, out);
    return hipGetLastError();
}

hipError_t launch_setup(const MpsInst& I, double* mslot, double* l0out, unsigned long long* dt0, hipStream_t st) {
    k_setup_elem<<<blocks_for(I.Ne, 256), 256, 0, st>>>(I, mslot, l0out, dt0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_gather_scalar<<<blocks_for(I.N, 256), 256, 0, st>>>(I, mslot, I.mass);
    return hipGetLastError();
}

hipError_t launch_gravity(const MpsInst& I, double grav, double* tslot, double* fgrav, hipStream_t st) {
    if (I.nen == 8) {
        k_grav_elem<8><<<blocks_for(I.Ne, 256), 256, 0, st>>>(I, grav, tslot);
        k_grav_node<8><<<blocks_for(I.N, 256), 256, 0, st>>>(I, tslot, fgrav);
    } else {
        k_grav_elem<4><<<blocks_for(I.Ne, 256), 256, 0, st>>>(I, grav, tslot);
        k_grav_node<4><<<blocks_for(I.N, 256), 256, 0, st>>>(I, tslot, fgrav);
    }
    return hipGetLastError();
}

hipError_t launch_fext(const MpsInst& I, double t, double* fext, hipStream_t st) {
    if (I.n_sgp > 0) k_press<<<blocks_for(I.n_sgp, 256), 256, 0, st>>>(I, 1, t);
    if (I.n_lnode > 0) k_loads<<<blocks_for(I.n_lnode, 256), 256, 0, st>>>(I, 1, t);
    k_fext<<<blocks_for(I.N, 256), 256, 0, st>>>(I, fext);
    return hipGetLastError();
}

hipError_t launch_loads(const MpsInst& I, hipStream_t st) {
    if (I.n_sgp > 0) k_press<<<blocks_for(I.n_sgp, 256), 256, 0, st>>>(I, 0, 0.0);
    if (I.n_lnode > 0) k_loads<<<blocks_for(I.n_lnode, 256), 256, 0, st>>>(I, 0, 0.0);
    return hipGetLastError();
}

hipError_t launch_fint(const MpsInst& I, double* fint, hipStream_t st) {
    k_gather_fint<<<blocks_for(I.N, 256), 256, 0, st>>>(I, fint);
    return hipGetLastError();
}

hipError_t launch_elem_outputs(const MpsInst& I, double* out, hipStream_t st) {
    k_elem_outputs<<<blocks_for(I.Ne, 256), 256, 0, st>>>(I, out);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t st) {
    k_fill_u64<<<(unsigned)std::min<long long>((n + 255) / 256, 1024), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_init_ctl(const MpsInst& I, const unsigned long long* dt0, hipStream_t st) {
    k_init_ctl<<<1, 64, 0, st>>>(I, dt0);
    return hipGetLastError();
}

hipError_t launch_energy(const MpsInst& I, unsigned flags, hipStream_t st) {
    const unsigned nb = (unsigned)I.n_patch;
    if (nb == 0) return hipSuccess;
    switch (I.law) {  // one instantiation per law: a runtime choice would size registers for every law
        case 0: k_energy<0><<<nb, MPS_PATCH, 0, st>>>(I, flags); break;
        case 1: k_energy<1><<<nb, MPS_PATCH, 0, st>>>(I, flags); break;
        case 2: k_energy<2><<<nb, MPS_PATCH, 0, st>>>(I, flags); break;
        default: k_energy<3><<<nb, MPS_PATCH, 0, st>>>(I, flags); break;
    }
    return hipGetLastError();
}

hipError_t launch_series(const MpsInst& I, hipStream_t st, unsigned nflags) {
    if (I.n_series <= 0) return hipSuccess;
    k_series<<<blocks_for(I.n_series, 256), 256, 0, st>>>(I, nflags);
    return hipGetLastError();
}

hipError_t launch_energy_tail(const MpsInst& I, hipStream_t st) {
    k_energy_tail<<<1, 256, 0, st>>>(I);
    return hipGetLastError();
}

template <int BS>
static hipError_t batch_phase_bs(int phase, int prec, int nen, int law, const MpsInst* insts, const MpsWork* work,
                                 const int* nwork, int grid, int fuse, unsigned nflags, hipStream_t st) {
    if (phase == 0) {
#define MPS_EB(T, NEN, LAW) k_elem_batch<T, NEN, LAW, BS><<<grid, BS, 0, st>>>(insts, work, nwork)
#define MPS_EB_LAW(T, NEN)                          \
    switch (law) {                                  \
        case 0: MPS_EB(T, NEN, 0); break;           \
        case 1: MPS_EB(T, NEN, 1); break;           \
        case 2: MPS_EB(T, NEN, 2); break;           \
        default: MPS_EB(T, NEN, 3); break;          \
    }
        if (prec == 1) {
            if (nen == 8) { MPS_EB_LAW(float, 8) } else { MPS_EB_LAW(float, 4) }
        } else {
            if (nen == 8) { MPS_EB_LAW(double, 8) } else { MPS_EB_LAW(double, 4) }
        }
#undef MPS_EB_LAW
#undef MPS_EB
    } else if (phase == 1) {
        k_node_batch<BS><<<grid, BS, 0, st>>>(insts, work, nwork, fuse, nflags);
    } else {
        k_begin_batch<BS><<<grid, BS, 0, st>>>(insts, work, nwork);
    }
    return hipGetLastError();
}

hipError_t launch_batch_phase(int phase, int bs, int prec, int nen, int law, const MpsInst* insts, const MpsWork* work,
                              const int* nwork, int grid, int fuse, unsigned nflags, hipStream_t st) {
    if (grid <= 0) return hipSuccess;
    switch (bs) {
        case 64: return batch_phase_bs<64>(phase, prec, nen, law, insts, work, nwork, grid, fuse, nflags, st);
        case 128: return batch_phase_bs<128>(phase, prec, nen, law, insts, work, nwork, grid, fuse, nflags, st);
        case 256: return batch_phase_bs<256>(phase, prec, nen, law, insts, work, nwork, grid, fuse, nflags, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mpsk
