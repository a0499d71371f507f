/*
 * mps_gpu.h -- C-ABI of the MI355X explicit-dynamics engine (libmps_gpu.so).
 *
 * Drop-in for the explicit hex8/tet4 neo-Hookean path of muphisim v1.1.  The
 * boundary is plain C: fixed-width integers, double pointers, SoA arrays and
 * 0-based indices.  Host buffers are read only during the call (copy-in); the
 * library owns all device memory.  No function calls exit(): every entry point
 * returns an int status (MPS_OK or a negative MPS_ERR_*), and mps_last_error()
 * returns the message of the last failure on a handle.
 *
 * Conventions kept from the reference:
 *   - DOF index = node + n_nodes*dir          (classGPs.cpp:54, classNodes.cpp:62)
 *   - nodal vectors are SoA by direction: u[dir*n_nodes + node]
 *   - element connectivity is element-major, reference node order
 *     (classLinearHexahedra.cpp:346-353; classLinearTetrahedra.cpp:270-292)
 *
 * A handle must not be used from two threads at once.  The server object
 * (mps_server_*) is internally synchronised.
 */
#ifndef MPS_GPU_H
#define MPS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MPS_OK 0
#define MPS_ERR_ARG (-1)      /* invalid argument or inconsistent input */
#define MPS_ERR_HIP (-2)      /* HIP runtime failure (no device, OOM, launch) */
#define MPS_ERR_DT (-3)       /* stable dt < 1e-20: "Negative jacobian" (solverExplicit.cpp:270-273) */
#define MPS_ERR_NAN (-4)      /* non-finite u[1] (solverExplicit.cpp:306-308) */
#define MPS_ERR_STATE (-5)    /* call not valid in the handle's current state */
#define MPS_ERR_CAPACITY (-6) /* request exceeds a configured capacity */

/* ---- run status reported in mps_report.status ---- */
#define MPS_RUNNING 0
#define MPS_DONE 1

/* ---- constitutive laws (constitutiveList.h; inputFile.cpp:1008-1014) ---- */
#define MPS_LAW_NH_COMPRESSIBLE 0 /* "*HyperElastic, 3D-Compressible-Neo-Hookean" */
#define MPS_LAW_NH_BELYTSCHKO 1   /* "*HyperElastic, Neo-Hookean" */
#define MPS_LAW_SVK 2             /* "*HyperElastic, St-Venant-Kirchhoff" (inputFile.cpp:1016-1020) */
#define MPS_LAW_LINEAR_ELASTIC 3  /* "*Linear-Elastic" (inputFile.cpp:886-928) */

/* ---- Dirichlet modes ---- */
#define MPS_BC_RAMP 0     /* *BOUNDARY, TYPE=DISPLACEMENT RAMP (classNodes.cpp:66-74) */
#define MPS_BC_INST 1     /* *BOUNDARY, TYPE=DISPLACEMENT INST (classNodes.cpp:58-64) */
#define MPS_BC_VELOCITY 2 /* extension: prescribed velocity V*min(step/ramp,1) */
#define MPS_BC_RAMP_RATE 3 /* ramp whose value is already UU = U/(tf-t0), as classDirichlet
                              stores it (classNodes.cpp:66-74, getUU classNodes.h:75-77) */

/* ---- precision forms ---- */
#define MPS_PREC_FP64 0      /* element math and everything else in FP64 */
#define MPS_PREC_FP32_ELEM 1 /* element math FP32; u/v/a, dt and energies FP64 */

typedef struct mps_solver_s *mps_handle;

typedef struct mps_mesh {
    int32_t n_nodes;
    int32_t n_elem;
    int32_t nen; /* 8 (C3D8, 2x2x2 Gauss) or 4 (C3D4, 1 Gauss point) */
    int32_t _pad;
    const double *X;     /* [3*n_nodes] reference coordinates, SoA */
    const int32_t *conn; /* [nen*n_elem] element-major, 0-based */
} mps_mesh;

typedef struct mps_material {
    int32_t law;         /* MPS_LAW_* */
    int32_t per_element; /* 0: E[0], nu[0], rho[0]; 1: arrays of n_elem */
    const double *E;
    const double *nu;
    const double *rho;
} mps_material;

typedef struct mps_bcs {
    int32_t n_dirichlet;
    int32_t mode;           /* MPS_BC_* (one mode per solver, as the reference) */
    int32_t vel_ramp_steps; /* MPS_BC_VELOCITY: steps to reach full amplitude */
    int32_t _pad;
    const int32_t *dof;  /* [n_dirichlet] node + n_nodes*dir, no duplicates */
    const double *value; /* [n_dirichlet] U (ramp/inst) or V (velocity) */
} mps_bcs;

/* ---- Neumann loads: one block per entry of the reference's NeumannBCs list, in input order ---- */
#define MPS_LOAD_FORCE 0         /* *BOUNDARY, TYPE=FORCE (classNodalForces; forceBCManagement, inputFile.cpp:2841-2941) */
#define MPS_LOAD_GRAVITY 1       /* *GRAVITY (classGravity; inputFile.cpp:4723-4735) */
#define MPS_LOAD_PRESSURE_RAMP 2 /* *BOUNDARY, TYPE=PRESSURE RAMP (classPressureRamp; inputFile.cpp:3580-3684) */
#define MPS_LOAD_PRESSURE_INST 3 /* *BOUNDARY, TYPE=PRESSURE INST (classPressureInst; inputFile.cpp:3686-3790) */
#define MPS_LOAD_TRACTION_RAMP 4 /* *BOUNDARY, TYPE=TRACTION RAMP (classTractionRamp; inputFile.cpp:3342-3460) */
#define MPS_LOAD_TRACTION_INST 5 /* *BOUNDARY, TYPE=TRACTION INST (classTractionInst; inputFile.cpp:3462-3578) */

/* The external force of every step is zero plus the blocks in order (NeumannBCManagement,
 * boundaryConditions.cpp:75-140):
 *   FORCE:    entry k adds force[3k..]/(tf-t0) * t at node index[k] for t0 <= t <= tf, else 0
 *             (BCsNodalForces :1120-1134; classNodes.h:101-108); force = F * Ff as read.
 *   GRAVITY:  f_z += -w rho |J| g phi_a per bulk Gauss point, g = value (BCsGravity :1146-1172).
 *   PRESSURE: face face[k] (getSubElement numbering: hex 1..6, tet 1..4) of element index[k] carries
 *             P(t) w |J_s| phi_a det(F) F^-T n0 per surface Gauss point, F of the bulk element at the point
 *             (BCsPressureRamp :370-452, BCsPressureInst :454-496); P(t) = value/(tf-t0) * (t-t0) (ramp) or
 *             value (inst) for t0 <= t <= tf, else 0 (classNeumannBCs.h:144-176).
 *   TRACTION: faces as for PRESSURE carry the dead load w |J_s| phi_a T_i(t) over the face's nodes
 *             (BCsTractionRamp / Inst :218-271); T(t) = force[0..2]/(tf-t0) * (t-t0) (ramp) or force[0..2]
 *             (inst) for t0 <= t <= tf, else 0 (classNeumannBCs.h:48-99). */
typedef struct mps_load {
    int32_t type;         /* MPS_LOAD_* */
    int32_t n;            /* FORCE: node entries; PRESSURE, TRACTION: faces; GRAVITY: 0 */
    double t0, tf;        /* FORCE, PRESSURE, TRACTION: window (tf > t0; the INST types: tf >= t0) */
    double value;         /* GRAVITY: g; PRESSURE: P */
    const int32_t *index; /* FORCE: node [n]; PRESSURE, TRACTION: bulk element [n] (0-based) */
    const int32_t *face;  /* PRESSURE, TRACTION: face [n] */
    const double *force;  /* FORCE: [3n]; TRACTION: [3] the traction vector */
} mps_load;

typedef struct mps_solver_cfg {
    double t0, tf;       /* *Time t0, tf */
    double scale;        /* *Scale Factor */
    int64_t n_steps;     /* 0: reference time-window loop; >0: fixed step count (extension) */
    int32_t precision;   /* MPS_PREC_* */
    int32_t energy_every;/* 0: no energies; k: energies every k-th step */
    int32_t history_cap; /* steps kept in the device history ring (0: 4096) */
    int32_t device;      /* HIP device ordinal */
} mps_solver_cfg;

typedef struct mps_report {
    int64_t steps;     /* steps completed by this solver so far */
    double time;       /* timeRun after the last completed step */
    double dt;         /* dt used by the last completed step */
    int32_t status;    /* MPS_RUNNING, MPS_DONE or a negative MPS_ERR_* */
    int32_t nonfinite; /* 1 if any non-finite value reached u[1] */
    double device_ms;  /* device time spent in this call's steps */
} mps_report;

/* Build a solver: uploads mesh, material and BCs, computes the GP geometry,
 * lumped mass and initial stable dt on the device.
 * Replaces: classSolverExplicit ctor + calculate() set-up
 * (solverExplicit.cpp:33-44, 63-200; computeLumpedMassVector :429-459;
 *  calculateCriticalTimeStep(nod,elem) solvers.cpp:276-287). */
int mps_create(const mps_mesh *mesh, const mps_material *mat, const mps_bcs *bcs,
               const mps_solver_cfg *cfg, mps_handle *out);

/* Hand over u, v, a from a previous solver (any pointer may be NULL = zeros).
 * Valid only before the first step.
 * Replaces: solvers::initValues (solvers.h:171-187) + setLastSolutionsToSystem
 * (solverExplicit.cpp:577-596). */
int mps_set_state(mps_handle h, const double *u, const double *v, const double *a);

/* Set the Neumann loads: n_loads blocks in input order (replacing any earlier call; 0 blocks: none).  At most
 * one GRAVITY block.  Valid only before the first step.  Every step evaluates them at its own time and
 * displacement: nodal forces and gravity in the node kernel, pressure faces in a surface kernel before it.
 * Replaces: NeumannCommonManagement's FORCE, GRAVITY and PRESSURE RAMP / INST entries
 * (inputFile.cpp:4703-4735) and their per-step evaluation in NeumannBCManagement (boundaryConditions.cpp:75-140). */
int mps_set_loads(mps_handle h, int32_t n_loads, const mps_load *loads);

/* External force fext[3*n_nodes] at time t for the solver's current displacement, from the loads of
 * mps_set_loads (device kernels; does not touch the solver state).
 * Replaces: NeumannBCManagement (boundaryConditions.cpp:75-140). */
int mps_external_force(mps_handle h, double t, double *fext);

/* ---- artificial viscosity (*Bulk viscosity / *Dev viscosity of a *SOLVER block) ---- */
#define MPS_VISC_NONE 0
#define MPS_VISC_BULK 1 /* bulkViscosity (solverExplicit.cpp:505-542); reader defaults c1 = 0.06, cl = 1.2 */
#define MPS_VISC_DEV 2  /* artificialViscosity (:462-503): its deviatoric stress never reaches the PK1, so it
                           adds no force; accepted and run as MPS_VISC_NONE */

/* Artificial viscosity of the solver (default none).  Valid only before the first step.  With MPS_VISC_BULK
 * every Gauss point adds the PK1 q cof(F), q = rho/J h (c1^2 h |trD| min(0, trD) + cl c0 trD),
 * trD = tr(F - F_prev)/dt, F_prev the point's F at the previous step (the state handed to the solver at
 * its first step).  The reference never initialises its element size and c0 (solvers.h:53-54): here they
 * are elem_size and c0 when > 0, else each element's undeformed characteristic length and sound speed.
 * mps_internal_force returns the elastic part only.
 * Replaces: the viscosity pass of computeInternalForce (solverExplicit.cpp:381-421). */
int mps_set_viscosity(mps_handle h, int32_t type, double c1, double cl, double elem_size, double c0);

/* Copy the current u, v, a to host (any pointer may be NULL).
 * Replaces: solvers::getFinalValues (solvers.h:171-187). */
int mps_get_state(mps_handle h, double *u, double *v, double *a);

/* Gather n entries of u (which=0), v (1) or a (2) at DOF indices
 * dof[i] = node + n_nodes*dir into out[i].
 * Replaces: the nodal value look-up of classExtractData::writeDataFromSolution
 * (dataExtraction.cpp:784-808), done on the device so that per-step nodal
 * extraction does not copy the whole field. */
int mps_get_dofs(mps_handle h, int32_t which, int64_t n, const int32_t *dof, double *out);

/* mps_get_dofs without any device allocation in the call: the indices and gathered values are staged in a
 * fixed per-device array of MPS_DOFS_STATIC_MAX entries, set up when the first solver is created on the
 * device and kept until the process ends.  n > MPS_DOFS_STATIC_MAX returns MPS_ERR_CAPACITY.  All calls on
 * one device share that array: calls that overlap in time on one device (from two threads) race.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (tight per-step
 * extraction or serving loops, harnesses that forbid hipMalloc).  It synchronises the solver's stream, so it
 * is not for use under stream capture. */
#define MPS_DOFS_STATIC_MAX 65536
int mps_get_dofs_static(mps_handle h, int32_t which, int64_t n, const int32_t *dof, double *out);

/* Internal force at n DOFs dof[i] = node + n_nodes*dir for the solver's current displacement, into out[i]:
 * the force the last step assembled (with its viscous part), summed from the step's partial nodal forces in
 * the order the update used; before the first step, the force at the initial displacement.  Only the listed
 * DOFs cross to the host.
 * Replaces: the intForce look-up of printForcesOut (output.cpp:153-279; intForce of computeRightHandSide,
 * solverExplicit.cpp:293-304) and of the InternalForce node extraction (dataExtraction.cpp:784-811). */
int mps_get_internal_force_dofs(mps_handle h, int64_t n, const int32_t *dof, double *out);

/* Advance up to max_steps explicit steps (fewer if the time window or the
 * fixed step count ends, or on error).  rep may be NULL.
 * Replaces: the time loop of classSolverExplicit::calculate
 * (solverExplicit.cpp:217-338). */
int mps_run(mps_handle h, int64_t max_steps, mps_report *rep);

/* Kinds of per-step DOF series (mps_set_dof_series). */
#define MPS_SERIES_U 0    /* displacement at the end of the step */
#define MPS_SERIES_V 1    /* velocity at the end of the step */
#define MPS_SERIES_A 2    /* acceleration at the end of the step */
#define MPS_SERIES_FINT 3 /* internal force of the step (the node kernel's sum of its partial forces) */
#define MPS_SERIES_FEXT 4 /* external force of the step (its loads at the step's time) */

/* Per-step series of single DOFs, recorded on the device inside every step (graph replay kept, no host round
 * trip per step) into a ring of history_cap rows: entry i is kind which[i] at DOF dof[i] = node +
 * n_nodes*dir.  Replaces any earlier series; n = 0 removes them.  mps_get_dof_series(h, first, count, out)
 * copies rows [first, first+count) (row = step index - 1, as mps_get_history) into out[count*n], row-major;
 * the step times are the history's time column.  Run in chunks of at most history_cap steps to read every
 * row.  Replaces: the per-step nodal look-up of classExtractData (Unknown, InternalForce, ExternalForce;
 * dataExtraction.cpp:778-912) and the internal force of printForcesOut (output.cpp:106-135), which the
 * reference reads from host vectors every step. */
int mps_set_dof_series(mps_handle h, int64_t n, const int32_t *which, const int32_t *dof);
int mps_get_dof_series(mps_handle h, int64_t first, int64_t count, double *out);

/* Per-step history from the device ring: entries [first, first+count) with
 * step index = entry + 1.  Energies are NaN on steps where they were not
 * computed.  Any output pointer may be NULL.  Energies cost one FP64 pass over the elements on each energy step
 * (in the element kernel itself for the FP64 form without viscosity) and keep graph replay and the fused
 * update; the KIN_ENERGY of a step is computed from its velocity when the next step starts (or at the end of the
 * run call), so it is complete once mps_run returns.
 * Replaces: classExtractData DEFO_ENERGY / KIN_ENERGY / VOLUME with the Sum
 * operation (dataExtraction.cpp:569-776). */
int mps_get_history(mps_handle h, int64_t first, int64_t count, double *time, double *dt,
                    double *e_def, double *e_kin, double *volume);

/* Energy balance of the same steps (solvers created with energy_every > 0; any pointer may be NULL), FP64:
 *   e_ext        EXTERNAL_ENERGY = sum_i u_i fext_i of the step (dataExtraction.cpp:867-875), on energy steps;
 *   e_kin_lumped 1/2 sum_i m_i v_i^2, the kinetic energy of the lumped-mass update, on energy steps;
 *   w_ext        work of the external loads since the end of step 1, trapezoidal over each step's increment
 *                du: sum 1/2 (fext_prev + fext) . du, every step (0 at step 1);
 *   w_react      work of the supports likewise: sum over constrained DOFs of 1/2 (r_prev + r) . du with
 *                r = fint - fext, plus the kinetic energy 1/2 m (v^2 - v_prev^2) the prescribed velocity gives the
 *                DOF's lumped mass (the update sets a = 0 there), every step (0 at step 1).
 * Balance: e_kin_lumped + DEFO_ENERGY - w_ext - w_react stays at its step-1 value up to the O(dt^2) error of
 * the explicit update.  No reference counterpart (the reference writes EXTERNAL_ENERGY only). */
int mps_get_energy_balance(mps_handle h, int64_t first, int64_t count, double *e_ext, double *e_kin_lumped,
                           double *w_ext, double *w_react);

/* Element fields of the current state for the VTK output: out[21*n_elem], per element the Gauss-point
 * averages of Cauchy stress (9, row-major), Green-Lagrange strain (9; the small strain for the linear
 * law), von Mises stress, volumetric strain and equivalent strain, law by law as the reference's
 * updateConstitutive sets them (0 where a law does not).
 * Replaces: the per-GP state accumulation of vtkPlotting2D3D (output.cpp:567-625) behind the
 * TENSORS Cauchy / E and SCALARS VMS / VolumetricStrain / EquivalentStrain sections. */
int mps_element_outputs(mps_handle h, double *out);

/* Output snapshots without stopping the solve.  mps_snapshot_begin(h, slot, with_elements) stages
 * u, v, a of the last completed step (and, with with_elements != 0, the 21 element output fields of
 * mps_element_outputs) on the device in stream order and starts their copy to pinned host memory on
 * a second stream; it returns at once, and mps_run may be called again right away.
 * mps_snapshot_wait(h, slot, &u, &v, &a, &elem) blocks until that copy is done and returns pointers
 * to it (3*n_nodes doubles each, SoA as mps_get_state; elem 21*n_elem as mps_element_outputs, or
 * NULL), valid until mps_snapshot_release(h, slot); a slot is begun again only after its release.
 * Two slots (0, 1) let one snapshot be written out while the next is taken.  mps_snapshot_wait and
 * mps_snapshot_release may run on another thread than the one calling mps_run; one slot's calls must
 * not overlap.  The first begin on a slot allocates its buffers (device and pinned host,
 * (9*n_nodes + 21*n_elem)*8 bytes each), kept until mps_destroy.
 * Replaces: the synchronous copy of the solution before writeVTK / vtkPlotting2D3D at each output
 * step (solverExplicit.cpp:310-322, output.cpp:296-918). */
int mps_snapshot_begin(mps_handle h, int32_t slot, int32_t with_elements);
int mps_snapshot_wait(mps_handle h, int32_t slot, const double **u, const double **v, const double **a,
                      const double **elem);
int mps_snapshot_release(mps_handle h, int32_t slot);

/* Lumped mass per node (identical for x, y, z).
 * Replaces: classSolverExplicit::computeLumpedMassVector (solverExplicit.cpp:429-459). */
int mps_get_mass(mps_handle h, double *mass);

/* Internal force fint[3*n_nodes] for displacement u (device kernels; does not
 * touch the solver state).
 * Replaces: classSolverExplicit::computeInternalForce (solverExplicit.cpp:363-379). */
int mps_internal_force(mps_handle h, const double *u, double *fint);

/* Stable dt s*min_e min(l0,l(u))/c for displacement u (NULL: undeformed).
 * Replaces: solvers::calculateCriticalTimeStep (solvers.cpp:276-320). */
int mps_critical_dt(mps_handle h, const double *u, double *dt);

/* Release the device memory of a solver.  Replaces: `delete nlsys` and the
 * solver destructor (solverExplicit.cpp:360; main.cpp:183-185). */
int mps_destroy(mps_handle h);

/* Graph replay of step chunks (default on): when enabled and kernel timing is off, mps_run
 * replays each chunk of steps from a captured hipGraph (one per chunk length and, with energies,
 * per first step modulo energy_every) instead of launching its kernels one by one.  No reference
 * counterpart. */
int mps_set_graphs(mps_handle h, int32_t enable);

/* Capture and instantiate, without running anything, the graphs that mps_run(h, max_steps) would
 * replay from the current step, so that the capture cost does not fall inside a later timed call.
 * Does nothing when mps_run would launch the kernels one by one.  No reference counterpart. */
int mps_prepare_run(mps_handle h, int64_t max_steps);

/* Kernel timing: when enabled, mps_run records HIP events around every kernel
 * launch on the solver's stream.  mps_kernel_times returns, per kernel kind
 * [0 element, 1 node, 2 begin, 3 energy], the summed device milliseconds and the
 * launch count since the last mps_profile call.
 * Replaces: the wall/CPU timers of the reference (solvers.cpp:22-32,
 * solverExplicit.cpp:358) with per-kernel device times. */
int mps_profile(mps_handle h, int32_t enable);
int mps_kernel_times(mps_handle h, double *ms_total /*[4]*/, int64_t *launches /*[4]*/);

/* Storage layout of h (no reference counterpart; the element and node kernels' bytes follow from it):
 * the number of patches (MPS_PATCH-element work units of the element kernel), of partial forces (the
 * distinct nodes of each patch, summed over the patches), and, when elem_order is not NULL, the input id
 * of each stored element (elem_order[n_elem]).  Hex8 elements are stored sorted along a space-filling
 * curve of their centroids, so that each patch is a compact block (MPS_ELEM_ORDER=input in the environment
 * at mps_create keeps the input order); tet4 elements in input order.  Every other entry point takes and
 * returns elements in input order. */
int mps_layout(mps_handle h, int64_t *n_patch, int64_t *n_part, int32_t *elem_order);

/* Message of the last failure on h (h may be NULL for mps_create failures). */
const char *mps_last_error(mps_handle h);

/* Name and properties of the HIP device used (no reference counterpart). */
int mps_device_info(int32_t device, char *buf, int32_t len);

/* Measured device-memory copy bandwidth (read + write bytes per second, GB/s) of a
 * grid-stride copy of `bytes` bytes, averaged over `reps` launches.  No reference
 * counterpart: it gives the attainable HBM ceiling for the roofline and is the
 * calibration kernel for hardware traffic counters. */
int mps_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double *copy_gbs);

/* ---------------------------------------------------------------------- */
/* Batched form: many independent solves packed into one device arena and  */
/* stepped together by the same kernels (no reference counterpart: the     */
/* reference runs one solve per process).                                  */
/* ---------------------------------------------------------------------- */

typedef struct mps_batch_s *mps_batch;

typedef struct mps_batch_cfg {
    int64_t arena_bytes;  /* device memory reserved for instance state (0: 1 GiB) */
    int32_t max_instances;/* instance slots (0: 1024) */
    int32_t precision;    /* MPS_PREC_* for every instance */
    int32_t steps_per_launch; /* steps captured per graph launch (0: 32) */
    int32_t elem_block;   /* work-items per element workgroup, 64/128/256 (0: 256 until mps_batch_autotune;
                             the server tunes it per mix size when both this and grid_mult are 0) */
    int32_t grid_mult;    /* persistent grid = grid_mult * CUs workgroups (0: 2 until mps_batch_autotune;
                             tuned by the server as elem_block) */
    int32_t device;
} mps_batch_cfg;

int mps_batch_create(const mps_batch_cfg *cfg, mps_batch *out);
/* Admit one solve; returns its slot id in *slot.  MPS_ERR_CAPACITY when the
 * arena or slot table is full.  The solve must use fixed n_steps (> 0). */
int mps_batch_add(mps_batch b, const mps_mesh *mesh, const mps_material *mat,
                  const mps_bcs *bcs, const mps_solver_cfg *cfg, int32_t *slot);
/* Advance every active solve by max_steps (rounded up to whole graph launches of
 * steps_per_launch steps; solves stop by themselves at their own n_steps) and wait. */
int mps_batch_step(mps_batch b, int64_t max_steps);
/* Enqueue `launches` graph launches without waiting (continuous batching). */
int mps_batch_launch(mps_batch b, int64_t launches);
/* Wait for enqueued launches and refresh the per-slot progress read by mps_batch_query. */
int mps_batch_sync(mps_batch b);
/* Arena occupancy: free bytes, largest free block, active solves (any pointer may be NULL). */
int mps_batch_info(mps_batch b, int64_t *free_bytes, int64_t *largest_block, int32_t *active);
/* Progress of one slot. */
int mps_batch_query(mps_batch b, int32_t slot, mps_report *rep);
/* Copy a slot's state out (any pointer NULL to skip). */
int mps_batch_get_state(mps_batch b, int32_t slot, double *u, double *v, double *a);
/* Free a slot for reuse. */
int mps_batch_remove(mps_batch b, int32_t slot);
/* Autotune elem_block / grid_mult on the current batch content; writes the choice. */
int mps_batch_autotune(mps_batch b, int32_t *elem_block, int32_t *grid_mult);
int mps_batch_destroy(mps_batch b);
const char *mps_batch_last_error(mps_batch b);

/* ---------------------------------------------------------------------- */
/* Serving form: a request loop with continuous batching over one batch     */
/* (no reference counterpart: the reference runs one solve per process).   */
/* Requests are admitted in arrival order when a slot and an arena block of */
/* their device footprint are free; their state stays on the device until  */
/* they finish.  The server object is internally synchronised.              */
/* ---------------------------------------------------------------------- */

#define MPS_REQ_QUEUED 0
#define MPS_REQ_RUNNING 1
#define MPS_REQ_DONE 2
#define MPS_REQ_FAILED 3

typedef struct mps_server_s *mps_server;

typedef struct mps_server_cfg {
    mps_batch_cfg batch; /* arena, slots, precision, steps per graph launch, tuning */
} mps_server_cfg;

typedef struct mps_server_stats {
    int64_t submitted, completed, failed, launches;
    double p50_ms, p99_ms, max_ms;  /* latency: submission -> result on the host */
    double queue_p50_ms;            /* submission -> admission */
    double span_ms;                 /* first submission -> last completion */
    double solves_per_s;            /* completed / span */
    double busy_ms;                 /* time with a graph launch in flight */
    int32_t elem_block, grid_mult;  /* tuning of the latest launch */
    int32_t tunings;                /* mix sizes the online tuner has settled */
    int32_t _pad;
} mps_server_stats;

int mps_server_create(const mps_server_cfg *cfg, mps_server *out);
/* Validate and prepare a request on the calling thread, then queue it; returns its id. */
int mps_server_submit(mps_server s, const mps_mesh *mesh, const mps_material *mat, const mps_bcs *bcs,
                      const mps_solver_cfg *cfg, int64_t *id);
/* Non-blocking: *state = MPS_REQ_*.  When finished, returns the solve's status code and, if u
 * is not NULL, copies the final displacement [3*n_nodes] and releases the request; lat_ms
 * (may be NULL) receives {latency, queue wait} in milliseconds. */
int mps_server_poll(mps_server s, int64_t id, int32_t *state, double *u, double *lat_ms);
/* Block until the request finishes or timeout_s elapses. */
int mps_server_wait(mps_server s, int64_t id, double timeout_s, int32_t *state);
/* Block until every submitted request finished (MPS_ERR_STATE on timeout). */
int mps_server_drain(mps_server s, double timeout_s);
int mps_server_get_stats(mps_server s, mps_server_stats *out);
int mps_server_reset_stats(mps_server s);
int mps_server_destroy(mps_server s);
const char *mps_server_last_error(mps_server s);

#ifdef __cplusplus
}
#endif
#endif
