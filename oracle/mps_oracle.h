/*
 * mps_oracle.h -- CPU oracle for the explicit hex8/tet4 neo-Hookean path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it restates the
 * reference algorithm (muphisim v1.1, classSolverExplicit and its callees)
 * in plain C so that tests, the smoke check and the CPU baseline leg of the
 * benchmark can compare the HIP path against it.  The product library never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line whose operation order it
 * follows.  Pins (what checks each function beyond retyping it) live in
 * tests/test_oracle_pins.py.
 *
 * Layout conventions (same as the reference):
 *   - DOF index = node + n_nodes * dir            (classGPs.cpp:54, classNodes.cpp:62)
 *   - X, u, v, a are SoA by direction: X[dir*n_nodes + node]
 *   - conn is element-major: conn[e*nen + a], 0-based node ids
 *   - 3x3 tensors are row-major [i*3+j]            (maths.h:331-338)
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LAW_NH_COMPRESSIBLE = 0, ORC_LAW_NH_BELYTSCHKO = 1, ORC_LAW_SVK = 2, ORC_LAW_LINEAR = 3 };
enum { ORC_BC_RAMP = 0, ORC_BC_INST = 1, ORC_BC_VELOCITY = 2 };
enum { ORC_SUM_REFERENCE = 0, ORC_SUM_ELEMENT = 1 };
enum { ORC_OK = 0, ORC_ERR_DT = -1, ORC_ERR_NAN = -2, ORC_ERR_ARG = -3, ORC_ERR_MEM = -4 };

enum { ORC_VISC_NONE = 0, ORC_VISC_BULK = 1, ORC_VISC_DEV = 2 };
enum { ORC_LOAD_FORCE = 0, ORC_LOAD_GRAVITY = 1, ORC_LOAD_PRESSURE_RAMP = 2, ORC_LOAD_PRESSURE_INST = 3,
       ORC_LOAD_TRACTION_RAMP = 4, ORC_LOAD_TRACTION_INST = 5 };

/* One Neumann block.
 *   FORCE (forceBCManagement, inputFile.cpp:2841-2941): n nodes index[], forces force[3n] = F * Ff over [t0, tf];
 *     F(t) = force/(tf-t0) * t inside the window, else 0 (classNeumann, classNodes.h:101-108, classNodes.cpp:85-96).
 *   GRAVITY (inputFile.cpp:4723-4735): value = g; f_z += -w rho |J| g phi_a per GP (boundaryConditions.cpp:1146-1172).
 *   PRESSURE RAMP / INST (pressure*BCManagement, inputFile.cpp:3580-3790): n faces, bulk element index[] and
 *     face[] (getSubElement numbering); value = P; P(t) = P/(tf-t0) * (t-t0) (ramp) or P (inst) inside
 *     [t0, tf], else 0 (classNeumannBCs.h:144-176); Nanson force, boundaryConditions.cpp:370-496.
 *   TRACTION RAMP / INST (traction*BCManagement, inputFile.cpp:3342-3578): faces as for pressure, force[3] the
 *     traction T; T(t) = T/(tf-t0) * (t-t0) (ramp) or T (inst) inside [t0, tf], else 0 (classNeumannBCs.h:48-99);
 *     force += w |J_s| phi_a T_i over the face's nodes, boundaryConditions.cpp:218-271. */
typedef struct orc_load {
    int32_t type;          /* ORC_LOAD_* */
    int32_t n;
    double t0, tf, value;
    const int32_t *index;  /* FORCE: node; PRESSURE: bulk element (0-based) */
    const int32_t *face;   /* PRESSURE: face S, 1..6 (hex) or 1..4 (tet) */
    const double *force;   /* FORCE: [3n]; TRACTION: [3] */
} orc_load;

typedef struct orc_problem {
    int32_t n_nodes;
    int32_t n_elem;
    int32_t nen;            /* 8 (C3D8) or 4 (C3D4) */
    int32_t law;            /* ORC_LAW_* */
    const double *X;        /* [3*n_nodes] SoA */
    const int32_t *conn;    /* [nen*n_elem] element-major */
    int32_t mat_per_elem;   /* 0: E/nu/rho have length 1, 1: length n_elem */
    const double *E;
    const double *nu;
    const double *rho;
    int32_t n_dir;          /* Dirichlet records in input order */
    const int32_t *dir_dof; /* node + n_nodes*dir */
    const double *dir_val;  /* ramp/inst: prescribed U; velocity: amplitude V */
    int32_t bc_mode;        /* ORC_BC_* */
    int32_t vel_ramp_steps; /* velocity mode: V_n = V*min(n/ramp,1) */
    double t0, tf, scale;
    int64_t n_steps;        /* 0: reference time-window loop; >0: fixed step count */
    int32_t sum_order;      /* ORC_SUM_* */
    int32_t n_threads;      /* >1 requires ORC_SUM_ELEMENT */
    int32_t energy_every;   /* 0: no energies; k: every k-th step (and the last) */
    int32_t _pad;
    /* Neumann loads in input order, the reference's NeumannBCs list (NeumannCommonManagement,
     * inputFile.cpp:4703-4760); fext of rhs = -fint + fext is rebuilt from them every step
     * (NeumannBCManagement, boundaryConditions.cpp:75-140) */
    int32_t n_loads;
    int32_t _pad2;
    const struct orc_load *loads;
    /* Artificial viscosity (*Bulk viscosity / *Dev viscosity, inputFile.cpp:2462-2510; solverExplicit.cpp:381-421,
     * 469-542), reference summation order only.  The reference's element size and c0 (solvers.h:53-54) are never
     * set; here they are visc_h and visc_c0 when > 0, else the element's undeformed length l0 and sound speed. */
    int32_t visc;           /* ORC_VISC_* */
    int32_t _pad3;
    double visc_c1, visc_cl;
    double visc_h, visc_c0;
} orc_problem;

typedef struct orc_report {
    int64_t steps;          /* steps taken */
    double time;            /* final timeRun */
    double dt_last;         /* dt of the last step */
    int32_t status;         /* ORC_* */
    int32_t _pad;
} orc_report;

/* History buffers (any may be NULL).  Entry k describes step k+1. */
typedef struct orc_history {
    int64_t cap;
    double *time;
    double *dt;
    double *e_def;          /* DEFO_ENERGY, Sum       (dataExtraction.cpp:569-629) */
    double *e_kin;          /* KIN_ENERGY, Sum        (dataExtraction.cpp:631-707) */
    double *volume;         /* VOLUME, Sum            (dataExtraction.cpp:709-776) */
    /* Energy balance (no reference counterpart; any may be NULL).  On energy steps: */
    double *e_ext;          /* EXTERNAL_ENERGY, sum_i u_i fext_i (dataExtraction.cpp:867-875) */
    double *e_kin_lumped;   /* 1/2 sum_i m_i v_i^2, the kinetic energy of the lumped-mass update */
    /* On every step, trapezoidal work since the end of step 1 (0 at step 1), with du the step's increment:
     * w_ext   = sum over steps and DOFs of 1/2 (fext_prev + fext) . du
     * w_react = sum over steps and constrained DOFs of 1/2 (r_prev + r) . du, r = fint - fext the support force,
     *           plus the kinetic energy the prescribed velocity gives the DOF's lumped mass, 1/2 m (v^2 - v_prev^2)
     *           (the update sets a = 0 on constrained DOFs, so r carries no inertia) */
    double *w_ext;
    double *w_react;
} orc_history;

/* ---- component functions (exported for the pin tests) ---- */
double orc_det3(const double *A);
void orc_inverse3(const double *A, double *Ainv);
void orc_hex8_shape(const double *xw, double *phi, double *natdev);
void orc_tet4_shape(const double *xw, double *phi, double *natdev);
int32_t orc_quadrature(int32_t nen, double *xw);
void orc_define_gp(int32_t nen, const double *xyz, const double *xw,
                   double *phi, double *dphi, double *wJ, double *detJ);
double orc_area_triangle(const double *tri);
double orc_char_length(int32_t nen, const double *xyz);
void orc_material(double E, double nu, double *lambda, double *mu, double *K1);
double orc_sound_speed(double E, double nu, double rho);
/* Sound speed of any law: NH sqrt((lambda+2mu)/rho), St-Venant-Kirchhoff and linear elastic sqrt(E/rho). */
double orc_law_sound_speed(int32_t law, double E, double nu, double rho);
void orc_pk1(int32_t law, const double *F, double E, double nu, double *P);
double orc_defo_energy(int32_t law, const double *F, double E, double nu);
void orc_kinematics(int32_t nen, const double *dphi, const double *uloc, double *F);
void orc_gp_force(int32_t nen, double wJ, const double *P, const double *dphi, double *f);
void orc_gp_lumped_mass(int32_t nen, double wJ, double rho, const double *phi, double *m);

/* ---- whole explicit solve (solverExplicit.cpp:63-361) ----
 * u, v, a: [3*n_nodes] initial state in, final state out. */
int32_t orc_run(const orc_problem *p, double *u, double *v, double *a,
                orc_history *hist, orc_report *rep);

/* Lumped mass per node (same value for x, y, z); Dirichlet-ness ignored. */
int32_t orc_lumped_mass(const orc_problem *p, double *mass);

/* Internal force fint[3*n_nodes] for a given displacement. */
int32_t orc_internal_force(const orc_problem *p, const double *u, double *fint);

/* Internal force with the step's artificial viscosity: u now, uprev at the previous step, step dt. */
int32_t orc_internal_force_visc(const orc_problem *p, const double *u, const double *uprev, double dt, double *fint);

/* DEFO_ENERGY, KIN_ENERGY (consistent interpolation of v at the GPs) and VOLUME, Sum over all GPs, for u, v. */
int32_t orc_energies(const orc_problem *p, const double *u, const double *v, double *out);

/* Critical time step s*min_e min(l0,l(u))/c (solvers.cpp:296-320); u may be NULL. */
double orc_critical_dt(const orc_problem *p, const double *u);
/* Per element (out[21*n_elem]): GP averages of Cauchy (9), Green-Lagrange strain (9; small strain for
 * the linear law), von Mises stress, volumetric and equivalent strain, as the VTK writer prints them
 * (output.cpp:567-820). */
int32_t orc_element_outputs(const orc_problem *p, const double *u, double *out);
/* External force fext[3*n_nodes] at time t for displacement u (NULL: zero), from the loads. */
int32_t orc_external_force(const orc_problem *p, const double *u, double t, double *fext);

/* Surface Gauss points of one pressure face (classLinearSquare / classLinearTriangles::defineGPs with the
 * bulk shape at each point, defineShapeOnAPoint): returns the number of points (4 hex, 1 tet), or < 0.
 * Per point g: xyz[3g..] position, ws[g] = w |J_s|, n0[3g..] = getn0 of the face, phi_s[g*4..] the face
 * shape values, phi_b[g*8..] and dphi_b[g*24..] the bulk shape values and gradients at the point. */
int32_t orc_pressure_face(const orc_problem *p, int32_t elem, int32_t face, double *xyz, double *ws, double *n0,
                          double *phi_s, double *phi_b, double *dphi_b);

#ifdef __cplusplus
}
#endif
#endif
