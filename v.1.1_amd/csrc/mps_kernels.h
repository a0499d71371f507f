// mps_kernels.h -- host-side launchers of the HIP kernels (internal to libmps_gpu).
#pragma once
#include <hip/hip_runtime.h>

#include "mps_common.h"

namespace mpsk {
// mode 0: step (dt reduction, ctl protocol); mode 1: standalone force evaluation.  eflags (mode 0): energies
// of the step (MPS_E_DEF, MPS_E_KIN) computed by the element kernel itself when elem_fuses_energy(I).
hipError_t launch_elem(const MpsInst& I, int mode, hipStream_t st, unsigned eflags = 0u);
// The element kernels of I compute energies themselves (FP64 form, no viscosity); otherwise the standalone pass.
bool elem_fuses_energy(const MpsInst& I);
// fuse = 1: also begin step k+1 (prediction + BC) in the same pass.  nflags: where the previous step left its
// velocity state (MPS_N_IN_VA: v and a, the first step of a launch; MPS_N_IN_WA: the half-step velocity in a's
// array, after a step with MPS_N_OUT_V; otherwise the half-step velocity in v's array); MPS_N_OUT_V to keep v
// readable for the next element phase; with energies on, the element-phase energies of the step to add up
// (MPS_E_DEF, MPS_E_KIN) and MPS_N_PREV when the previous step ran in this launch.
hipError_t launch_node(const MpsInst& I, int fuse, hipStream_t st, unsigned nflags);
hipError_t launch_begin(const MpsInst& I, hipStream_t st);
hipError_t launch_dt(const MpsInst& I, const double* u, unsigned long long* out, hipStream_t st);
hipError_t launch_setup(const MpsInst& I, double* mslot, double* l0out, unsigned long long* dt0, hipStream_t st);
hipError_t launch_fint(const MpsInst& I, double* fint, hipStream_t st);
// Gravity load fgrav[N] (z) for g; tslot: scratch of Ne*ngp doubles.
hipError_t launch_gravity(const MpsInst& I, double grav, double* tslot, double* fgrav, hipStream_t st);
// External force fext[3N] at time t from the instance's loads (surface forces from the current u).
hipError_t launch_fext(const MpsInst& I, double t, double* fext, hipStream_t st);
// External force of the flagged nodes for the step in flight (surface kernel, then the per-node sums);
// runs after the step's element kernel, before its node kernel.
hipError_t launch_loads(const MpsInst& I, hipStream_t st);
// Standalone energy pass: MPS_E_DEF / MPS_E_KIN of step ctl->B between its element and node kernels, or
// MPS_E_KIN_NOW after the last node kernel of a launch; block sums only (added up by the next kernel).
hipError_t launch_energy(const MpsInst& I, unsigned flags, hipStream_t st);
// Per-step DOF series of step ctl->B into their ring (between the step's element and node kernels); nflags: those
// of the step's node kernel.
hipError_t launch_series(const MpsInst& I, hipStream_t st, unsigned nflags);
// End of a launch with energies, after the MPS_E_KIN_NOW pass: the last step's sums into the history.
hipError_t launch_energy_tail(const MpsInst& I, hipStream_t st);
// Element fields of the VTK output: out[21*Ne] (Cauchy 9, strain 9, VMS, volumetric, equivalent strain), rows in
// input element order.
hipError_t launch_elem_outputs(const MpsInst& I, double* out, hipStream_t st);
hipError_t launch_init_ctl(const MpsInst& I, const unsigned long long* dt0, hipStream_t st);
// Persistent batched phases: 0 element, 1 node (fuse flag, nflags as for launch_node), 2 begin.  bs in {64,128,256}.
// The batch must be homogeneous in precision, element type (nen) and law.
// nwork points to the item count in device memory (graphs stay valid when tables change).
hipError_t launch_batch_phase(int phase, int bs, int prec, int nen, int law, const MpsInst* insts, const MpsWork* work,
                              const int* nwork, int grid, int fuse, unsigned nflags, hipStream_t st);
// Fill n 64-bit words with v (device kernel: no memset node in captured graphs).
hipError_t launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t st);
}  // namespace mpsk
