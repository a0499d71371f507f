// mps_common.h -- structures shared by the host side of libmps_gpu and its HIP kernels.
//
// Device data layout (one solve = one "instance"):
//   X, u, v, a      double[3N] SoA by direction (DOF = node + N*dir, as the reference)
//   mass            double[N]    lumped mass, identical for the three directions
//   ninfo           uint32[N]    bits 0..2 Dirichlet mask (x,y,z), bits 3..30 Dirichlet record index,
//                                bit 31 node carries nodal force entries (MPS_NODE_LOADED)
//   bcval           double[3*R]  per Dirichlet record: per-direction BC value
//   conn            int32[nen*Ne] element-major (reference node order), elements in stored order (hex8:
//                                sorted along a space-filling curve of their centroids at set-up, for compact
//                                patches; tet4: input order); l0, mat, cel, rho and the element loops follow it
//   eorig           int32[Ne]    input id of each stored element (NULL: stored = input order)
//   l0              double[Ne]   undeformed characteristic length
//   mat             double[2*Ne] per-element (lambda,mu) or (mu,K1); NULL = uniform
//   cel             double[Ne]   per-element sound speed; NULL = uniform c0
//   rowptr/col      CSR node -> stored slot (e*nen + a), ascending input element then slot (set-up gathers)
//   patches         elements [p*MPS_PATCH, (p+1)*MPS_PATCH) form patch p; its U_p distinct
//                   nodes (ascending) own partial forces part[poff[p] .. poff[p+1])
//     poff          int32[n_patch+1]       first partial of each patch
//     prow          uint16[n_part+n_patch] per patch, U_p+1 offsets into its pcol list
//                                          (+8 entries of padding in the single-instance form)
//     pcol          uint16[nen*Ne]         per patch, contributions i*nen+a (local element i,
//                                          slot a) grouped by local node, ascending (+8 entries of
//                                          padding in the single-instance form)
//     nrow/ncol     CSR node -> partials, ascending patch (= ascending element)
//     pnd           int32[n_part]          tet4: node id of each partial (the patch's distinct nodes)
//     plx           uint16[nen*Ne]         tet4: local node index (within its patch) of each element slot
//   part            double[3*n_part]       partial nodal forces [(g)*3 + dir]; also the
//                                          set-up scratch (>= nen*Ne doubles)
//   loads           the n_lnode nodes with load items (flag bit 31 of ninfo), each with its items in input order:
//     lidx          int32[N]         compact index of a flagged node (read only for flagged nodes)
//     lnn, lrow     int32[n_lnode], [n_lnode+1]  node id; CSR into lcode
//     lcode         int32[items]     kind (bits 30-31: MPS_ITEM_*) | index
//     lf            double[3*n_lnode]  the step's external force of each flagged node
//     lm, lwin      double[3*n_f], [2*n_f]  nodal force entry: F/(tf-t0) (classNodes.cpp:85-96); t0, tf
//     fgrav         double[N]      gravity load along z (NULL: no gravity)
//     sge, sgb      int32[n_sgp]   pressure surface Gauss point: bulk element, block
//     sgd           double[MPS_SGP_STRIDE*n_sgp]  bulk dphi (24), n0 (3), w|J_s| (1), scatter phi (8)
//     sblk          double[8*blocks]  block type, P, t0, tf, traction (3), -
//     pc            double[24*n_sgp]  the step's surface forces [(g*8 + a)*3 + dir]
//   uprev           double[3N]   displacement of the previous step (bulk viscosity only; NULL otherwise)
//   ctl             MpsCtl       step control (time, dt, go flags), running work sums, reduction counters
//   dtred           u64[2*MPS_DT_SLOTS*MPS_DT_STRIDE]  stable-dt reduction slots by step parity; in the
//                   single-instance form followed by MPS_DT_SLOTS + 1 counters (same stride) of the
//                   stable-dt certificate (MPS_DT_COUNT_WORDS)
//   hist            double[MPS_HIST_COLS*hist_cap]     per-step history ring, column c at c*hist_cap
//   energy bookkeeping (solves with energies only):
//     epart         double[3*n_patch + 2*MPS_NODE_PARTS*node_blocks]  per-block energy sums: element phase, then
//                   the node kernel's by step parity
//     rprev         double[3*R]       previous step's support force fint - fext of each Dirichlet record
//     lfprev        double[3*n_lnode] previous step's external force of each flagged node
#pragma once
#include <stdint.h>

#define MPS_MAX_NEN 8
#define MPS_PATCH 256  // elements per patch (element-kernel work unit and partial-sum scope)
#define MPS_DT_SLOTS 16   // stable-dt reduction slots per step parity: workgroup b min-reduces into slot b % 16
#define MPS_DT_STRIDE 32  // u64 between two slots (256 B), so the slots' atomics do not share a line
#define MPS_DT_SKIP_HOLD 64  // steps without the stable-dt skip after a failed certificate
#define MPS_DT_COUNT_WORDS ((MPS_DT_SLOTS + 1) * MPS_DT_STRIDE)  // certificate counters after the dt slots
#define MPS_NODE_LOADED 0x80000000u  // ninfo bit: the node has nodal force entries
#define MPS_REC_MASK 0x7FFFFFFFu     // ninfo bits below the loaded flag
#define MPS_ITEM_FORCE 0u            // load item kinds (lcode >> 30)
#define MPS_ITEM_PRESSURE 1u
#define MPS_ITEM_GRAVITY 2u
#define MPS_ITEM_INDEX 0x3FFFFFFFu
#define MPS_SGP_STRIDE 36            // doubles per pressure surface Gauss point
#define MPS_SBLK_STRIDE 8            // doubles per surface-load block: type, P, t0, tf, Tx, Ty, Tz, -
// History ring columns: time, dt, DEFO_ENERGY, KIN_ENERGY, VOLUME, EXTERNAL_ENERGY, lumped kinetic energy,
// work of the loads, work of the supports
enum : int32_t { MPS_H_TIME = 0, MPS_H_DT, MPS_H_EDEF, MPS_H_EKIN, MPS_H_VOL, MPS_H_EEXT, MPS_H_EKL, MPS_H_WEXT,
                 MPS_H_WREACT, MPS_HIST_COLS };
#define MPS_NODE_PARTS 4  // node-kernel partial sums per block: work of loads, work of supports, lumped KE, u.fext
// Energy flags of an element-kernel launch (fused energies) and of the standalone energy pass
#define MPS_E_DEF 1u   // DEFO_ENERGY and VOLUME of the step in flight (u_k)
#define MPS_E_KIN 2u   // KIN_ENERGY of the previous step (v_{k-1}, before the solve of step k)
#define MPS_E_KIN_NOW 4u  // standalone pass after the last node kernel of a chunk: KIN_ENERGY of that step
#define MPS_N_PREV 8u  // node kernel: also reduce the previous step's node sums (that step ran in the same launch)
// node kernel: where the previous step left its velocity state (default: the half-step velocity w in v's array)
#define MPS_N_IN_VA 16u   // v and a of the previous step in their arrays
#define MPS_N_IN_WA 32u   // w in a's array (the previous step kept v readable: MPS_N_OUT_V)
#define MPS_N_OUT_V 64u   // node kernel: keep v of the step readable (kinetic energy) and put w in a's array

enum : int32_t { MPS_TIME_WINDOW = 0, MPS_TIME_FIXED = 1 };

// Per-step control record, double-buffered by step parity.
struct MpsStepSlot {
    double time;  // timeRun of the step (lagged update, solverExplicit.cpp:227-232)
    double dt;    // dt of the step (solverExplicit.cpp:267-268)
    int32_t go;   // 1: step was begun (prediction + BC applied), 0: not run
    int32_t _pad;
};

// Step protocol (all in device memory, no host round trip per step):
//   A = next step to run, written only by the node kernel, read by begin/element kernels;
//   B = step whose forces are in part, written only by the element kernel, read by node/energy.
// slot[k&1] holds step k; the dtred slots of parity k&1 receive the min over elements for
// step k (computed from u_{k-1}); the node kernel of step k resets them.
struct MpsCtl {
    MpsStepSlot slot[2];
    long long A;
    long long B;
    int32_t status;     // MPS_RUNNING / MPS_DONE / MPS_ERR_DT
    int32_t nonfinite;  // set when u[1] is non-finite
    long long nan_step; // first step with non-finite u[1] (LLONG_MAX if none)
    unsigned long long dt0min;  // min_e l0/c (setup)
    double w_ext, w_react;      // work of the loads and of the supports since the end of step 1
    long long wred;             // last step whose node-kernel energy sums were reduced
    // stable-dt skip (FP64 inviscid hex8 element kernel): elements whose lower bound of min(l, l0)/c exceeds
    // dt_skip skip the strict length; the kernel's last workgroup certifies the step's min against it
    double dt_skip;             // threshold for the next element kernel (+inf: every element is exact)
    long long skip_hold;        // no skipping before this step (set when a certificate failed)
};

// Largest node count whose byte offsets within one direction array (8 * node) fit 32 bits.
constexpr long long MPS_OFF32_NODES = 1ll << 29;

struct MpsInst {
    int32_t N, Ne, nen, law;
    int32_t precision, bc_mode, vel_ramp, time_mode;
    int64_t n_steps;
    double t0, tf, scale;
    double mat_a, mat_b, c0, rho0;  // uniform material: (lambda,mu) or (mu,K1), sound speed, density
    const double* X;
    const int32_t* conn;
    const int32_t* eorig;  // input id of each stored element (NULL: input order)
    const double* l0;
    const double* mat;
    const double* cel;
    const double* rho;
    double* u;
    double* v;
    double* a;
    double* mass;
    const uint32_t* ninfo;
    const double* bcval;
    const int32_t* rowptr;
    const int32_t* col;
    const int32_t* poff;
    const uint16_t* prow;
    const uint16_t* pcol;
    const int32_t* nrow;
    const int32_t* ncol;
    double* part;
    unsigned long long* dtred;  // dt reduction slots (bit patterns of non-negative doubles, integer min)
    MpsCtl* ctl;
    double* hist;      // [MPS_HIST_COLS*hist_cap] (MPS_H_* columns)
    int32_t hist_cap;
    int32_t energy_every;  // 0: no energies; k: energies every k-th step, work sums every step
    double* epart;     // energy block sums: element phase [3*n_patch], node kernel [2][MPS_NODE_PARTS*node blocks]
    double* rprev;     // [3*R] previous support force per Dirichlet record
    double* lfprev;    // [3*n_lnode] previous external force per flagged node
    int32_t dof1_node, dof1_dir;
    int32_t n_patch;
    int32_t n_part;
    int32_t n_lnode;  // nodes with load items
    int32_t n_sgp;    // pressure surface Gauss points
    const int32_t* lidx;
    const int32_t* lnn;
    const int32_t* lrow;
    const int32_t* lcode;
    double* lf;
    const double* lm;
    const double* lwin;
    const double* fgrav;
    const int32_t* sge;
    const int32_t* sgb;
    const double* sgd;
    const double* sblk;
    double* pc;
    // bulk artificial viscosity (solverExplicit.cpp:381-421, 505-542): active when uprev != NULL
    double visc_c1, visc_cl;   // quadratic and linear coefficients
    double visc_h, visc_c0;    // element size and sound speed (> 0), or <= 0: the element's l0 and c
    double* uprev;
    // tet4 node staging: the element kernel loads each patch's distinct nodes once into LDS
    const int32_t* pnd;
    const uint16_t* plx;
    int32_t pu_max;  // largest number of distinct nodes of a patch
    // 1: the hex8 element kernels gather nodes by 32-bit byte offsets (at most MPS_OFF32_NODES nodes)
    int32_t off32;
    // stable-dt skip: on/off, and the threshold's margin over the previous step's min (MpsCtl::dt_skip)
    int32_t dt_skip_on;
    double dt_skip_margin;
    // per-step DOF series (mps_set_dof_series): n_series entries, recorded into ser_ring[(k-1)%hist_cap][i]
    int32_t n_series;
    const int32_t* ser_dof;   // node + N*dir
    const int32_t* ser_kind;  // MPS_SERIES_*
    double* ser_ring;
};

// One unit of work for the persistent batched kernels: a contiguous range of
// elements (or nodes) of one instance.
struct MpsWork {
    int32_t inst;
    int32_t begin;
    int32_t end;
    int32_t first;  // 1 if this item holds index 0 of the instance (scalar updates)
};
