/* dgx.h -- C ABI of the MI355X (gfx950) explicit-DG residual and RK engine.
 *
 * The engine replaces the hot path of DGM's compressible Euler and
 * Navier-Stokes domains: the residual Domain::time_derivative and the
 * explicit stepper TimeInt::advance.  Every entry point below names the
 * reference interface it stands in for (paths relative to the reference
 * root).  Plain pointers and sizes only; no C++ or torch types cross this
 * boundary.  All functions return DGX_OK (0) or a negative error code;
 * dgx_last_error() gives the message of the calling thread's last failure.
 *
 * Conventions
 *  - A "domain" owns the mesh tables, the FP64 conserved state U, the RK
 *    registers and all scratch on one GPU, and one HIP stream.  Calls on
 *    different domains are independent and may run concurrently; calls on
 *    one domain must not overlap.
 *  - The conserved state is kept in the reference's modal coefficients
 *    (normalized Legendre, src/Basis.cpp:74-102), layout [field][elem][mode]
 *    with modes ordered as in Quad::inner_product / Hex::inner_product
 *    (src/Quad.cpp:632-662, src/Hex.cpp:703-747).  Nodal values at the
 *    q = p+2 Gauss-Lobatto-Legendre points use layout [field][elem][node],
 *    nodes ordered i*q+j (quad) or (i*q+j)*q+k (hex) (src/Quad.cpp:767-789,
 *    src/Hex.cpp:817-866).
 *  - Residual sign follows the reference: F = -dU/dt (src/Domain.hpp:237-243).
 */
#ifndef DGX_H
#define DGX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGX_VERSION "0.1.0"

enum dgx_status {
  DGX_OK = 0,
  DGX_ERR_ARGUMENT = -1,      /* bad argument or unsupported configuration */
  DGX_ERR_DEVICE = -2,        /* HIP runtime failure */
  DGX_ERR_NONREALIZABLE = -3, /* density or pressure fell below aepsM
                                 (src/Euler.cpp:479-488) */
  DGX_ERR_MEMORY = -4         /* allocation failed */
};

/* Equation types: the reference's eqntype numbers (src/Reo.hpp:33-61). */
enum dgx_eqntype {
  DGX_EULER = 1,            /* Reo::Euler (2D)            */
  DGX_NAVIER_STOKES = 2,    /* Reo::Navier_Stokes (2D)    */
  DGX_EULER3D = 8,          /* Reo::Euler3d               */
  DGX_NAVIER_STOKES3D = 9,  /* Reo::Navier_Stokes3d       */
  DGX_SHALLOW = 30          /* Reo::Shallow (2D, fields phi, uphi, vphi;
                               Lax-Friedrichs flux, FP64) */
};

/* Numerical flux (src/Euler.cpp:102-113). */
enum dgx_flux { DGX_FLUX_LAX_FRIEDRICHS = 0, DGX_FLUX_ROE = 1 };

/* Viscous face gradient of Navier-Stokes.  BR1 is the reference's
 * (src/Navier_Stokes3d.cpp:182-391); BR2 (hexes only) has no reference
 * counterpart: on each face the viscous flux uses the element gradient plus
 * br2_eta times the lifting of that face's jump alone, instead of the gradient
 * lifted over all faces. */
enum dgx_viscous { DGX_VISCOUS_BR1 = 0, DGX_VISCOUS_BR2 = 1 };

/* Arithmetic of the residual.  DGX_MIXED: hexes (Euler3d, Navier_Stokes3d,
 * p <= 8) run their sum-factorized contractions on FP32 MFMA; affine quads
 * (Euler, Shallow) hold their element tensors in LDS and contract them in
 * FP32.  The state, RK registers, face traces, metric terms, Riemann fluxes
 * and pointwise physics stay FP64. */
enum dgx_precision { DGX_FP64 = 0, DGX_MIXED = 1 };

/* Time integrators: the reference's inttype numbers
 * (src/Domain.cpp:564-628): 3 TVD-RK3, 4 RK4, 14 LSRK14. */
enum dgx_inttype { DGX_TVD = 3, DGX_RK4 = 4, DGX_LSRK14 = 14 };

/* Node layout selector for state transfers. */
enum dgx_layout { DGX_NODAL = 0, DGX_MODAL = 1 };

typedef struct dgx_domain *dgx_handle;

/* Boundary sides (Euler3d, Shallow): a .con row (eid, sid, -1, type) marks
 * side sid of element eid as a boundary instead of a link.  Slip wall as
 * Euler3d::WallBC (src/Euler3d.cpp:876-909) or Shallow::WallBC
 * (src/ShallowBC.hpp:45-62), imposed state as StateBC (src/BC.cpp:17-28);
 * the ghost values meet the interior through the same Riemann flux as an
 * interior face. */
enum dgx_bc { DGX_BC_WALL = 0, DGX_BC_STATE = 1, DGX_BC_FUNCTION = 2, DGX_BC_ISOTHERMAL = 3 };
/* DGX_BC_ISOTHERMAL (Navier_Stokes3d): Isothermal_WallBC
 * (src/Navier_Stokes3d.cpp:472-512).  As in the reference, the inviscid flux
 * sees the slip Euler3d::WallBC that Euler3d::make_BC makes of every "wall"
 * code (src/Euler3d.cpp:857-860); the viscous terms see the ghost (rho-, 0, 0,
 * 0, (wall_Eref - 1/2) rho- wall_T) and the ghost viscous flux -S, so the face
 * viscous flux is the interior one.  Navier_Stokes3d state boundaries take
 * bc_state in both, with the same -S viscous ghost (StateBC::apply_flux,
 * src/BC.cpp:31-40).  Navier_Stokes3d has no slip-wall viscous flux
 * (Euler3d::WallBC::apply_flux raises, src/Euler3d.cpp:911-913), so
 * DGX_BC_WALL is rejected there. */

/* A function of (x, y, z, t) with ncomp components, as the reference's
 * FuncBC and FuncSource read them from root.bct / root.src and evaluate them
 * with its expression parser (src/BC.hpp, src/Source.hpp:190-230,
 * src/Parser.y).  Each component is a postfix program: code[start[c]] up to
 * code[start[c+1]] for c < ncomp.  An instruction holds its opcode in the low
 * 8 bits; DGX_OP_CONST pushes consts[code >> 8].  Binary operators pop b, then
 * a, and push a op b; a program leaves exactly one value and needs at most
 * DGX_FN_STACK slots.  The library copies the programs at creation. */
enum dgx_op {
  DGX_OP_CONST = 0, DGX_OP_X = 1, DGX_OP_Y = 2, DGX_OP_Z = 3, DGX_OP_T = 4,
  DGX_OP_ADD = 5, DGX_OP_SUB = 6, DGX_OP_MUL = 7, DGX_OP_DIV = 8, DGX_OP_POW = 9, DGX_OP_NEG = 10,
  DGX_OP_SIN = 11, DGX_OP_COS = 12, DGX_OP_TAN = 13, DGX_OP_ASIN = 14, DGX_OP_ACOS = 15,
  DGX_OP_ATAN = 16, DGX_OP_SINH = 17, DGX_OP_COSH = 18, DGX_OP_TANH = 19, DGX_OP_EXP = 20,
  DGX_OP_LOG = 21, DGX_OP_LOG10 = 22, DGX_OP_SQRT = 23, DGX_OP_ABS = 24, DGX_OP_ATAN2 = 25
};
#define DGX_FN_STACK 16
typedef struct {
  int ncomp;
  const int *start;      /* [ncomp + 1] */
  const int *code;       /* [start[ncomp]] */
  int nconst;
  const double *consts;  /* [nconst] */
} dgx_function;

/* Mesh and physics description.  `vert` holds, per element, nsd rows of
 * vertex coordinates exactly as in a root.msh file ([elem][dim][vertex],
 * 4 vertices for quads, 8 for hexes).  `con` holds `ncon` rows
 * (eid, sid, leid, lsid) in root.con order; the first row naming a side pair
 * makes that side the master (src/Field.cpp:3478-3556). */
typedef struct {
  int nsd;          /* 2 or 3 */
  int ne;           /* number of elements */
  int p;            /* polynomial order (L = p+1 modes, q = p+2 GLL points) */
  int eqntype;      /* enum dgx_eqntype */
  int flux;         /* enum dgx_flux */
  int precision;    /* enum dgx_precision */
  double gamma;     /* ratio of specific heats (default 1.4) */
  double vis;       /* dynamic viscosity mu (Navier-Stokes) */
  double Pr;        /* Prandtl number (default 0.72) */
  const double *vert;
  int ncon;
  const int *con;
  int viscous;      /* enum dgx_viscous (Navier-Stokes; 0 = BR1) */
  double br2_eta;   /* BR2 penalty; 0 selects 6, the number of hex faces */
  double bc_state[5];  /* conservative state of DGX_BC_STATE boundaries (Shallow: the first 3) */
  const double *xnodes;  /* curved elements, FP64: hexes (Euler3d, p <= 8; Navier_Stokes3d with
                            DGX_VISCOUS_BR1, p <= 7) coordinates [3][ne][q^3]
                            of every element's GLL nodes, node (i*q + j)*q + k; quads (Euler,
                            Navier_Stokes, Shallow) [2][ne][q^2], node i*q + j; NULL: affine */
  int no_balance;   /* Shallow: 1 leaves out the balance term.  By default, as
                       compiled in the reference (src/Shallow.cpp:251-266), the
                       domain's first residual is stored and subtracted from it
                       and from every later residual */
  const dgx_function *bc_function;  /* Shallow: (phi, uphi, vphi) of DGX_BC_FUNCTION
                                       boundaries at the side's Gauss points and the
                                       stage's time (FuncBC); NULL: none */
  const dgx_function *source;       /* Shallow: source s(x, y, t) at the nodes, F -= s
                                       (FuncSource::apply, src/Source.cpp:11-24); NULL: none.
                                       With either function, quads take the curved-element
                                       path (nodal metrics from xnodes, or from the
                                       vertices' bilinear map when xnodes is NULL) */
  double wall_T;    /* Navier_Stokes3d DGX_BC_ISOTHERMAL: wall temperature Tw and */
  double wall_Eref; /* reference energy Eref (Isothermal_Wall tag { Tw Eref } in root.bct) */
  int bulk_forcing; /* 1: hold the bulk density, x-momentum and energy, as a run
                       directory with root.chn does (Channel::holdBulkMassFlow,
                       src/Channel.hpp:117-163, called at src/Navier_Stokes3d.cpp:160):
                       after the projection the domain means of the rho, rho u and
                       rho E residuals are subtracted from them.  Affine hexes,
                       DGX_FP64 */
  const int *porder; /* variable order, as root.ord gives it (Field::Field,
                       src/Field.cpp:2290-2320): the order of every element [ne], p
                       their largest; NULL: every element has order p.  Affine quads,
                       Euler (or Navier_Stokes with vis = 0, whose diffusion adds
                       exact zeros), DGX_FP64, fixed dt (no dgx_set_time_step).
                       The two sides of a link carry max(L_left, L_right) Gauss
                       points (Side::match, src/Side.cpp:106-117); the lift
                       interpolates them to the element's GLL points
                       (Quad::local_lift, src/Quad.cpp:1971-2038).  State layouts
                       keep slots of the largest order: element e's (p_e+1)^2
                       modes (or (p_e+2)^2 nodes) come first in its slot of
                       (p+1)^2 (or (p+2)^2), zeros after */
} dgx_domain_desc;

/* Build a domain: geometry, side links and flips, basis tables, device
 * buffers.  `stream` is a hipStream_t or NULL (the library creates one).
 * Replaces the Field/Element construction of Reo::Euler::Euler /
 * Euler3d::Euler3d (src/Euler.cpp:47-211, src/Euler3d.cpp:47-150). */
int dgx_domain_create(const dgx_domain_desc *desc, void *stream, dgx_handle *out);
int dgx_domain_destroy(dgx_handle h);

/* Sizes: nodes and modes per element, fields, elements. */
int dgx_sizes(dgx_handle h, int *qtot, int *nmodes, int *nfields, int *ne);

/* vField totals of the domain's state (vField::size, qtot, nmodes, ndof,
 * sdof, nsd; src/vField.hpp:330-376, bound in python/dgm.pyx:297-318): fields;
 * quadrature points, modes and side quadrature points summed over all fields
 * and elements; space dimensions.  ndof equals nmodes (modal bases).  Any
 * pointer may be NULL. */
int dgx_vfield_sizes(dgx_handle h, int *size, long long *qtot, long long *nmodes, long long *sdof, int *nsd);

/* Nodal coordinates [dim][elem][node] to host memory (Element::C,
 * src/Quad.cpp:993-1013, src/Hex.cpp:991-1039). */
int dgx_coordinates(dgx_handle h, double *xyz);

/* State transfers.  Nodal uploads are L2-projected into the modal space, as
 * the reference's first restart write does (src/vField.cpp:1102-1181).
 * `on_device` = 1 if the pointer is device memory. */
int dgx_state_upload(dgx_handle h, const double *U, int layout, int on_device);
int dgx_state_download(dgx_handle h, double *U, int layout, int on_device);

/* Device pointer of the modal state ([field][elem][mode], FP64). */
int dgx_state_device_ptr(dgx_handle h, double **U);

/* Residual F = -dU/dt of the current state into the domain's residual
 * buffer (Domain::time_derivative, src/Domain.hpp:237-243;
 * Euler::time_derivative src/Euler.cpp:282-287; Euler3d src/Euler3d.cpp:455;
 * Navier_Stokes(3d)::time_derivative src/Navier_Stokes.cpp:89-94,
 * src/Navier_Stokes3d.cpp:135-161). */
int dgx_rhs(dgx_handle h);

/* Solution time (Domain::t): the time at which dgx_rhs evaluates boundary and
 * source functions.  dgx_advance moves it through each stage as the
 * reference's integrators do (RK4: t + dt/2 after the first and third
 * stages; TVD-RK3: c = 0, 1, 1/2; LSRK14: t_n + c_s dt) and leaves t_n + dt. */
int dgx_set_time(dgx_handle h, double t);
int dgx_get_time(dgx_handle h, double *t);
int dgx_residual_download(dgx_handle h, double *F, int layout, int on_device);

/* Advance nsteps steps of size dt with integrator `inttype`
 * (TimeInt::advance: RungeKutta src/TimeInt.cpp:267-302, TVD :824-856,
 * LSRK14 :1989-2013).  Asynchronous on the domain's stream. */
int dgx_advance(dgx_handle h, int inttype, double dt, int nsteps);

/* CFL time step (Domain::set_time_step, called before every step by
 * Domain::solve, src/Domain.cpp:931).  For 2D domains (Euler, Navier_Stokes,
 * Shallow) as Euler::set_time_step (src/Euler.cpp:804-877) and
 * Shallow::set_time_step (src/Shallow.cpp:573-637) with DGM_STORE_METRICS:
 * from the current state at the GLL nodes,
 *   idt = (|u dxr + v dyr| + |u dxs + v dys| + c acoustic) ivol
 * with the metrics of Element::compute_metrics_2d_tensor_prod
 * (src/Element.cpp:485-555); *cfl = dt * max idt.  With CFLmax != 0 the new
 * step is CFLmax * dt / cfl (no larger than dtmax when dtmax != 0), else dt;
 * *cfl is then rescaled to the new step.  For hexes the reference's
 * Euler3d::set_time_step (src/Euler3d.cpp:462-465) keeps dt and returns it as
 * the CFL number, and so does this call.  Waits for the domain's stream;
 * DGX_ERR_NONREALIZABLE if the state gives no finite sound speed. */
int dgx_set_time_step(dgx_handle h, double dt, double CFLmax, double dtmax, double *cfl, double *dt_new);

/* dgx_set_time_step without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * kernels' scratch (one maximum per element plus the domain maximum) is a
 * file-scope device array of 2^19 doubles (4 MiB, held by every GPU the
 * library runs on until the process ends), so domains of more than 524,287
 * elements are refused with DGX_ERR_ARGUMENT.  Same kernels and results as
 * dgx_set_time_step.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race. */
int dgx_set_time_step_noalloc(dgx_handle h, double dt, double CFLmax, double dtmax, double *cfl, double *dt_new);

/* One stage of a step: stage `stage` (0 .. 3 RK4, 0 .. 2 TVD-RK3, 0 .. 13
 * LSRK14) of integrator `inttype`, exactly as dgx_advance runs it, so a
 * device-resident TimeInt can do host work between stages (the per-stage body
 * of RungeKutta::advance src/TimeInt.cpp:267-302, TVD::advance :824-856,
 * LSRK14::advance :1989-2013: residual at the stage time, then the register
 * update).  Stages must come in order from 0, with one integrator and dt, and
 * the last one completes the step; dgx_advance is refused while a step begun
 * here is unfinished.  Asynchronous on the domain's stream. */
int dgx_stage(dgx_handle h, int inttype, int stage, double dt);

/* Wait for the domain's stream; returns DGX_ERR_NONREALIZABLE if any
 * residual met a non-realizable state since the last check. */
int dgx_synchronize(dgx_handle h);

/* Per-field Linf, L2, H1 norms of the state, out[3*field + {0,1,2}]
 * (Field::norm, src/Field.cpp:4675-4697; Quad::norm_L2/H1 src/Quad.cpp:2556-2632). */
int dgx_norms(dgx_handle h, double *out);

/* Number of RHS evaluations (stages) performed so far. */
long long dgx_stage_count(dgx_handle h);

/* Diagnostic: the most conjugate-gradient iterations any curved element's
 * mass-matrix solve has taken since the domain was created (0 on affine
 * domains and on curved quads, which use Cholesky factors).  The reference
 * factors each curved element's mass matrix instead (src/Hex.cpp:289-318). */
int dgx_solver_iterations(dgx_handle h, int *max_iters);

/* Run dgx_advance with HIP events around every kernel on the domain's
 * stream and return the summed milliseconds (`ms`) and launch counts per
 * kernel kind: 0 face traces, 1 residual, 2 lifted gradients, 3 viscous
 * fluxes, 4 nodal->modal, 5 modal->nodal, 6 stage updates.  `nkinds` >= 7.
 * Plays the role of the reference's FunctionTimer instrumentation
 * (src/Utilities.hpp:185-212) for the device path. */
int dgx_profile(dgx_handle h, int inttype, double dt, int nsteps, double *ms, int *count,
                int nkinds);

/* Device bytes a domain built from `desc` holds (state, RK registers,
 * traces, viscous terms, nodal transfer buffer, geometry), without building
 * it.  Used for admission against a memory budget. */
int dgx_domain_bytes(const dgx_domain_desc *desc, size_t *bytes);

/* ---------------------------------------------------------------------------
 * Serving loop: continuous batching of independent forward solves.
 *
 * No counterpart in the reference.  It drives the step seam (dgx_advance) for
 * optimisation and UQ ensembles.
 *
 * Admission and overlap:
 *  - Requests are admitted FIFO as slots (max_active) and budget (hbm_budget,
 *    by dgx_domain_bytes) free up.
 *  - Each admitted request gets its own domain and HIP stream, so the work of
 *    different requests overlaps on the device.
 *
 * Stepping:
 *  - Each pass of dgx_server_progress enqueues the next slice of
 *    steps_per_slice steps of every admitted request that has none in flight.
 *  - The first slice of a request runs directly.
 *  - The second slice is captured into a hipGraph, which later full slices
 *    replay, so small meshes are not bound by launch overhead.
 *  - With `batch`, 2D Euler and shallow-water LSRK14 requests of one order
 *    and precision form a group: each stage is one residual-and-update
 *    launch over the elements of every member, which also writes the next
 *    stage's face traces (a face-trace launch runs only at a slice's first
 *    stage), on a fixed grid of workgroups that walk the members' elements.
 *    Members step in lockstep slices and join or leave at slice boundaries.
 *  - A request's result does not depend on what else runs: each element's
 *    arithmetic sees only its own request's buffers.
 *
 * Not thread safe: one host thread drives a server. */
typedef struct dgx_server *dgx_server_handle;

typedef struct {
  size_t hbm_budget;     /* device bytes the admitted requests may hold together */
  int max_active;        /* admitted requests at once */
  int steps_per_slice;   /* steps enqueued per request and pass */
  int use_graphs;        /* replay slices from a captured hipGraph (1) or launch directly (0) */
  int batch;             /* 1: batch 2D Euler / shallow-water LSRK14 requests of one order and
                            precision into shared launches (one per stage for all of them) */
  int persistent;        /* with batch: 1 runs each group slice as one persistent cooperative
                            kernel (every stage, grid-wide barriers between stages) instead of
                            two launches per stage; results are bitwise the same */
} dgx_server_config;

enum dgx_request_state { DGX_QUEUED = 0, DGX_RUNNING = 1, DGX_DONE = 2, DGX_FAILED = 3 };

int dgx_server_create(const dgx_server_config *cfg, dgx_server_handle *out);
int dgx_server_destroy(dgx_server_handle s);

/* Queue a forward solve of `nsteps` steps of size `dt` with integrator
 * `inttype` from the initial state U0 (host memory, `layout` as in
 * dgx_state_upload).  `desc` and U0 are copied.  Fails with DGX_ERR_ARGUMENT
 * if the request alone exceeds the budget. */
int dgx_server_submit(dgx_server_handle s, const dgx_domain_desc *desc, const double *U0, int layout,
                      int inttype, double dt, int nsteps, long long *id);

/* One scheduling pass (non-blocking): retire finished slices and requests,
 * admit queued requests, enqueue slices.  *pending = requests not yet done or
 * failed. */
int dgx_server_progress(dgx_server_handle s, int *pending);

/* State of a request (enum dgx_request_state). */
int dgx_server_status(dgx_server_handle s, long long id, int *state);

/* Result of a finished request: the final state in `layout` (may be NULL),
 * and times[3] = seconds from server creation at submission, admission and
 * completion (may be NULL). */
int dgx_server_result(dgx_server_handle s, long long id, double *U, int layout, double *times);

/* Peak device bytes held by admitted requests, and peak admitted count. */
int dgx_server_peak(dgx_server_handle s, size_t *bytes, int *active);

/* ---------------------------------------------------------------------------
 * Acoustic inverse-source problem and its discrete adjoint.
 *
 * The linearized 2D Euler equations about a uniform mean flow, in conservative
 * perturbation variables: the reference's LinEuler_quasi (eqntype 4,
 * src/LinEuler_quasi.cpp:151-495) with a uniform mean, Lax-Friedrichs flux.
 * LSRK14 from a zero state, driven by theta_e g(t) in the energy equation
 * (theta constant per element, g a Ricker wavelet).  After every step the
 * pressure perturbation p' = (gamma-1)(E' - u0 m' - v0 n' + |u0|^2 rho'/2) is
 * observed at the receivers; J = 1/2 dt sum_steps sum_r (p'_r - d_r)^2.
 * dJ/dtheta comes from the discrete adjoint of the whole time loop (the
 * reversed LSRK14 recurrences, cf. LSRK14::adjoint_advance,
 * src/TimeInt.cpp:2064-2092, with the transposed DG residual), the forward
 * states being recomputed from binomial checkpoints (revolve; cf.
 * DTK::Griewank, src/DTK_Griewank.hpp:42-125).  FP64 only.
 * ------------------------------------------------------------------------- */
enum { DGX_LINEULER = 4 };   /* eqntype of the mesh description */

typedef struct dgx_acoustic *dgx_acoustic_handle;

typedef struct {
  double rho0, u0, v0, p0;   /* uniform mean flow */
  int nrec;                  /* number of receivers */
  const double *xrec;        /* receiver points [nrec][2] */
  double f0, t0;             /* g(t) = (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2 */
  double dt;                 /* time step */
  int nsteps;                /* LSRK14 steps */
  int nsnap;                 /* checkpoint slots, state 0 included (0: ceil(sqrt(nsteps)) + 1) */
} dgx_acoustic_desc;

/* `mesh`: nsd 2, eqntype DGX_LINEULER, precision DGX_FP64 or DGX_MIXED (FP32
 * element contractions in the forward and adjoint kernels, FP64 state,
 * registers, observations and gradient accumulation; flux and vis are
 * ignored).  Allocates nsnap + 4 states of 4 x ne x (p+1)^2 doubles. */
int dgx_acoustic_create(const dgx_domain_desc *mesh, const dgx_acoustic_desc *ad, void *stream,
                        dgx_acoustic_handle *out);
int dgx_acoustic_destroy(dgx_acoustic_handle a);

/* Forward solve with source theta[ne]; y[nsteps][nrec] receives the
 * observations (host memory, may be NULL). */
int dgx_acoustic_forward(dgx_acoustic_handle a, const double *theta, double *y);

/* Objective J and gradient grad[ne] = dJ/dtheta for data[nsteps][nrec]. */
int dgx_acoustic_gradient(dgx_acoustic_handle a, const double *theta, const double *data, double *J,
                          double *grad);

/* Forward and adjoint steps taken so far, and the most checkpoint slots the
 * last gradient held at once. */
int dgx_acoustic_stats(dgx_acoustic_handle a, long long *forward_steps, long long *adjoint_steps,
                       int *snapshots);

/* The checkpointing schedule alone (host only): forward steps it advances to
 * reverse nsteps steps with nsnap slots, the most slots it holds, and whether
 * every step was reversed in order from its own forward state. */
int dgx_revolve_count(int nsteps, int nsnap, long long *advances, int *max_slots, int *order_ok);

/* ---------------------------------------------------------------------------
 * Discrete adjoint of the nonlinear 2D Euler equations.
 *
 * LSRK14 steps of the 2D Euler residual (Euler::convection with the
 * Lax-Friedrichs flux, lift and projection, src/Euler.cpp:296-757,
 * src/Quad.cpp:1971-2038) from an initial modal state U0 [4][ne][(p+1)^2].
 * After every step the pressure (gamma-1)(E - |m|^2 / (2 rho)) is observed at
 * the receivers; J = 1/2 dt sum_steps sum_r (p_r - d_r)^2.  dJ/dU0 comes from
 * the reversed LSRK14 recurrences (LSRK14::adjoint_advance,
 * src/TimeInt.cpp:2064-2092) with the transposed residual Jacobian at every
 * stage state (hand-written backward kernels of the residual), the forward
 * states being recomputed from binomial checkpoints (cf. DTK::Griewank,
 * src/DTK_Griewank.hpp:42-125).  Periodic affine quads, FP64.
 * ------------------------------------------------------------------------- */
typedef struct dgx_euler_adjoint *dgx_euler_adjoint_handle;

typedef struct {
  int nrec;                  /* number of receivers */
  const double *xrec;        /* receiver points [nrec][2] */
  double dt;                 /* time step */
  int nsteps;                /* LSRK14 steps */
  int nsnap;                 /* checkpoint slots, state 0 included (0: ceil(sqrt(nsteps)) + 1) */
} dgx_euler_adjoint_desc;

/* `mesh`: nsd 2, eqntype DGX_EULER, vis 0, DGX_FLUX_LAX_FRIEDRICHS or
 * DGX_FLUX_ROE (the backward kernels apply the chosen flux's transposed
 * Jacobian; the Roe flux's is assembled from forward-mode derivatives of the
 * same formula), DGX_FP64, periodic parallelogram quads (no boundary rows,
 * xnodes, functions or porder).  Allocates nsnap + 20 states of
 * 4 x ne x (p+1)^2 doubles. */
int dgx_euler_adjoint_create(const dgx_domain_desc *mesh, const dgx_euler_adjoint_desc *ad, void *stream,
                             dgx_euler_adjoint_handle *out);
int dgx_euler_adjoint_destroy(dgx_euler_adjoint_handle a);

/* Forward solve from U0 (host); y[nsteps][nrec] receives the observations and
 * UT the final modal state (host memory, each may be NULL).  Returns
 * DGX_ERR_NONREALIZABLE if a density or pressure fell below aepsM. */
int dgx_euler_adjoint_forward(dgx_euler_adjoint_handle a, const double *U0, double *y, double *UT);

/* Objective J and gradient grad = dJ/dU0 [4][ne][(p+1)^2] for data[nsteps][nrec]. */
int dgx_euler_adjoint_gradient(dgx_euler_adjoint_handle a, const double *U0, const double *data, double *J,
                               double *grad);

/* One transposed residual Jacobian: out = (dR/dU(U))^T lam, R = -dU/dt the
 * modal residual of dgx_rhs (host arrays [4][ne][(p+1)^2]). */
int dgx_euler_adjoint_vjp(dgx_euler_adjoint_handle a, const double *U, const double *lam, double *out);

/* Forward and adjoint steps taken so far, and the most checkpoint slots the
 * last gradient held at once. */
int dgx_euler_adjoint_stats(dgx_euler_adjoint_handle a, long long *forward_steps, long long *adjoint_steps,
                            int *snapshots);

/* ---------------------------------------------------------------------------
 * Element-block seam: the batched hex transforms of GpGpu::hex_kernels
 * (src/GpGpu.hpp:24-57, src/GpGpu.cu:174-420 and 470-800), which
 * HexBlock::initialize_gpu / forward_transform / backward_transform drive for
 * a block of hexes that share one basis (src/HexBlock.cpp:35-130).
 *
 * Every element owns a slot of qa*qb*qc doubles, [elem][a][b][c] with c
 * fastest; its nmodes^3 modal coefficients are the first values of the slot,
 * [i][j][k] with k fastest.  Both transforms work in place, as hex_kernels'
 * do, on host memory (copied in and out on the block's stream, as the
 * reference does on every call) or on device memory (`on_device` = 1: no
 * copies; the library runs on the caller's buffer).
 * ------------------------------------------------------------------------- */
typedef struct dgx_hexblock *dgx_hexblock_handle;

/* hex_kernels::hex_kernels + initialize(ne, nmodes, qa, qb, qc): sizes
 * (1 <= nmodes <= qa, qb, qc <= 16) and the block's device buffers.
 * `stream` is a hipStream_t or NULL (the library creates one). */
int dgx_hexblock_create(long long ne, int nmodes, int qa, int qb, int qc, void *stream, dgx_hexblock_handle *out);
/* hex_kernels::~hex_kernels */
int dgx_hexblock_destroy(dgx_hexblock_handle h);

/* hex_kernels::wJ: the weighted Jacobian wJ [qa][qb][qc] (host memory), one
 * for the whole block as HexBlock::initialize_gpu passes it (affine hexes),
 * or with per_element = 1 one per element [ne][qa][qb][qc]. */
int dgx_hexblock_set_wJ(dgx_hexblock_handle h, const double *wJ, int per_element);

/* hex_kernels::Basis: Ba [nmodes][qa], Bb [nmodes][qb], Bc [nmodes][qc]
 * (host memory, row-major, the reference's psi->Ba/Bb/Bc). */
int dgx_hexblock_set_basis(dgx_hexblock_handle h, const double *Ba, const double *Bb, const double *Bc);

/* hex_kernels::inner_product (Hex::inner_product, src/Hex.cpp:703-747):
 * f *= wJ, then the a, b and c contractions with Ba, Bb, Bc; each slot's
 * first nmodes^3 values become the inner products and the rest of the slot
 * is left as it was.  Synchronous for host memory, asynchronous on the
 * block's stream for device memory. */
int dgx_hexblock_inner_product(dgx_hexblock_handle h, double *u, int on_device);

/* hex_kernels::backward_transform (Hex::backward_transform,
 * src/Hex.cpp:796-812): the first nmodes^3 values of each slot (modal) to the
 * slot's qa*qb*qc values at the points.  Synchronous for host memory,
 * asynchronous on the block's stream for device memory. */
int dgx_hexblock_backward_transform(dgx_hexblock_handle h, double *u, int on_device);

/* Wait for the block's stream. */
int dgx_hexblock_synchronize(dgx_hexblock_handle h);

const char *dgx_last_error(void);
const char *dgx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DGX_H */
