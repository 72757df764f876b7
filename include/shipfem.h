/*
 * shipfem.h -- C ABI of the MI355X explicit tetrahedral FEM step.
 *
 * The drop-in boundary for the reference's Vulkan tet solver
 * (engine/gpu/fem_gpu.hpp:17-45, class gpu::FemGpuSolver) and the CPU set-up
 * functions it consumes (engine/sim/fem.hpp:38-86). Plain C: opaque handles,
 * plain pointers and sizes, status codes, no exceptions, no torch types.
 *
 * Units: SI. Node arrays are xyz-interleaved per node ([3N]) at the boundary, as
 * in sim::fem::TetMesh (engine/sim/fem.hpp:38-57); the library stores SoA planes on
 * the device. Tet connectivity is [4T] node ids.
 *
 * Threading: a solver, batch or mesh handle is used by one host thread at a time
 * (the reference solver is single-threaded and blocking, fem_gpu.cpp:117-125).
 * A server handle owns its own submission thread; its functions are thread-safe.
 *
 * Errors: every entry point returns a shipfem_status. SHIPFEM_NO_DEVICE is the
 * "degrade, not abort" case of fem_gpu.hpp:21-22: callers skip, they do not fail.
 */
#ifndef SHIPFEM_H
#define SHIPFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHIPFEM_ABI_VERSION 1

typedef enum shipfem_status {
    SHIPFEM_OK = 0,
    SHIPFEM_NO_DEVICE = 1,      /* no usable gfx950 device: skip, do not fail */
    SHIPFEM_INVALID_ARG = 2,    /* refused input, message says which */
    SHIPFEM_OUT_OF_MEMORY = 3,  /* device or host allocation failed */
    SHIPFEM_UNSTABLE = 4,       /* a case produced a non-finite value */
    SHIPFEM_DEVICE_ERROR = 5,   /* HIP runtime error */
    SHIPFEM_BUSY = 6,           /* server: footprint exceeds the admission budget */
    SHIPFEM_NOT_FOUND = 7,
    SHIPFEM_UNSUPPORTED = 8
} shipfem_status;

const char* shipfem_status_string(shipfem_status s);

/* Node flags (one uint32 per node). PIN_* hold a DOF at zero velocity, as the
 * reference's per-node `fixed` (fem.cpp:197-200) does for all three; DRIVE_*
 * prescribe a smoothstep-ramped displacement (zone.cpp:1080-1145); LOADED marks a
 * node with a nonzero external load; CONTACT makes it a candidate for the indenter. */
#define SHIPFEM_PIN_X 0x01u
#define SHIPFEM_PIN_Y 0x02u
#define SHIPFEM_PIN_Z 0x04u
#define SHIPFEM_PIN_ALL 0x07u
#define SHIPFEM_LOADED 0x08u
#define SHIPFEM_CONTACT 0x10u
#define SHIPFEM_DRIVE_X 0x20u
#define SHIPFEM_DRIVE_Y 0x40u
#define SHIPFEM_DRIVE_Z 0x80u

/* sim::fem::Material (engine/sim/fem.hpp:24-36) plus the linear isotropic J2 flow
 * curve of sim::plasticity (plasticity.hpp:110-128, Hardening::Linear). */
typedef struct shipfem_material {
    double youngs_modulus;    /* Pa */
    double poisson_ratio;
    double density;           /* kg/m^3 */
    int32_t plastic;          /* 0: co-rotational linear elastic; 1: + J2 */
    double yield_strength;    /* Pa, sigma_y(0) */
    double hardening_modulus; /* Pa, H of sigma_y = sigma_y0 + H eps_p */
    /* Ductile failure and element deletion (plastic only; engine/sim/plasticity.cpp:457-477,
     * plasticity.hpp:160-210). failure_strain <= 0 disables it. Damage grows by
     * d eqps / (failure_strain * exp(-sensitivity (eta - reference_triaxiality))), with eta
     * the triaxiality of the returned stress and none at or below cutoff_triaxiality; a tet
     * whose damage reaches 1 tears: it carries no stress from then on. The reference's
     * constants are 1.5, 1/3 and -1/3; see shipfem_regularised_failure_strain. */
    double failure_strain;
    double triaxiality_sensitivity;
    double reference_triaxiality;
    double cutoff_triaxiality;
} shipfem_material;

/* Rigid sphere driven at constant velocity, node-to-surface penalty contact:
 * f = penalty * m_i / dt^2 * gap * n for candidate nodes inside the sphere. */
typedef struct shipfem_indenter {
    int32_t enabled;
    double radius;      /* m */
    double center[3];   /* m, at t = 0 */
    double velocity[3]; /* m/s */
    double penalty;     /* dimensionless kappa; 0 => 0.1 */
} shipfem_indenter;

/* One case: mesh, material, boundary conditions, loads, contact and time step.
 * Arrays are borrowed only for the duration of the create call. */
typedef struct shipfem_case_desc {
    uint32_t node_count, tet_count;
    const double* rest_position;        /* [3N] rest coordinates X */
    const uint32_t* index;              /* [4T] node ids per tet */
    const uint32_t* node_flags;         /* [N] SHIPFEM_* or NULL */
    const double* load;                 /* [3N] nodal force at full ramp, N; or NULL */
    const double* prescribed;           /* [3N] displacement of DRIVE_* DOF; or NULL */
    const double* initial_displacement; /* [3N] u at step 0 (x = X + u); or NULL */
    const double* initial_velocity;     /* [3N] v_{-1/2}; or NULL */
    shipfem_material material;
    shipfem_indenter indenter;
    double load_ramp;    /* s, load scale min(t/load_ramp, 1); 0 => full at once */
    double drive_ramp;   /* s, smoothstep ramp of the prescribed drive; 0 => immediate */
    double gravity;      /* m/s^2 along -z (fem.cpp:201) */
    double damping;      /* velocity multiplier per step (fem.cpp:209-210); 0 => 1 */
    double dt;           /* s; 0 => dt_safety * 2 / omega_max (FP64, on the device) */
    double dt_safety;    /* 0 => 0.9 */
    int32_t sample_every; /* energy-record period in steps, even; 0 => 100 */
    int64_t max_steps;   /* a case stops advancing at this step; 0 => unbounded */
    /* [N] lumped nodal mass, kg, in place of density * volume / corners per element corner
     * (the caller-owned TetMesh::mass of fem.hpp:42); or NULL. A node that is not fully
     * pinned and lies in an element needs a positive mass. The default stable step is then
     * the density bound times sqrt(min over such nodes of mass / (density share)), which
     * bounds the lumped system's largest frequency from below. */
    const double* mass;
    /* Solid-shell hexes with ductile failure: the local fracture strain eps_f
     * (Failure::fractureStrain, plasticity.hpp:194-195). When > 0, every element resolves
     * its own failure strain from its size, as initialisePlasticState does
     * (solid_shell.cpp:808-815): material.failure_strain is then the uniform strain eps_g,
     * and the element takes eps_g + (eps_f - eps_g) min(1, thickness / in-plane size)
     * (shipfem_hex_failure_strain). 0: every element takes material.failure_strain. Tet cases
     * take material.failure_strain as given and refuse a nonzero value. */
    double fracture_strain;
} shipfem_case_desc;

/* Per-call step parameters of gpu::FemGpuSolver::run (fem_gpu.hpp:31-32). */
typedef struct shipfem_step_params {
    double dt;
    double gravity;
    double damping;
    double youngs_modulus, poisson_ratio; /* <= 0 keeps the case's */
} shipfem_step_params;

/* Energy account (FP64), modelled on zone.cpp:1014-1070. At a record for step n:
 * strain = U(u_n), kinetic = 1/2 sum m v_{n-1/2} v_{n+1/2}, works integrated to
 * t_n by the trapezoid rule, dissipation = sum V sigma : d eps_p through step n.
 * balance residual = w_ext + w_contact + w_boundary - (strain + kinetic - E_0)
 *                    - dissipation - damping_loss. */
typedef struct shipfem_energy_record {
    int64_t step;
    double time;
    double strain, kinetic, dissipation, w_ext, w_contact, w_boundary, damping_loss;
    double fracture;         /* stored energy released by torn tets, 1/2 V sigma:(eps - eps_p) */
    double contact_force[3]; /* sum of penalty forces on the structure at step n */
} shipfem_energy_record;

/* Readback destinations; any pointer may be NULL. */
typedef struct shipfem_state_view {
    double* displacement;  /* [3N] u (FP32 + compensation, widened) */
    double* velocity;      /* [3N] v_{n-1/2} */
    float* position;       /* [3N] X + u in float (the reference's readback, fem_gpu.cpp:424-428) */
    float* eqps;           /* [T] equivalent plastic strain */
    float* plastic_strain; /* [6T] Voigt, engineering shears, co-rotated frame */
    float* damage;         /* [T] ductile damage, 1 = torn (0 without failure) */
    int64_t step;          /* out */
    double time;           /* out */
    uint32_t status;       /* out: nonzero when the case went non-finite */
} shipfem_state_view;

/* ---- device -------------------------------------------------------------- */
/* Name and HBM size of device 0; SHIPFEM_NO_DEVICE when there is none. */
shipfem_status shipfem_device_query(char* name, size_t name_len, uint64_t* hbm_bytes);

/* ---- meshes and case building (host; engine/sim/fem.cpp:119-263) ---------- */
typedef struct shipfem_mesh shipfem_mesh;

/* Kuhn box, reference numbering and tet order (makeBoxTetMesh, fem.cpp:217-263). */
shipfem_status shipfem_mesh_box(double lx, double ly, double lz, int32_t nx, int32_t ny,
                                int32_t nz, shipfem_mesh** out);
/* Box mesh jitter: amp * spacing * U(-1,1) per axis from a counter-based seeded
 * generator, except components normal to a boundary face. Refuses a jitter that
 * inverts a tet. Only valid on meshes made by shipfem_mesh_box. */
shipfem_status shipfem_mesh_jitter(shipfem_mesh* m, double amp, uint64_t seed);
/* Stiffened panel: plate lx*ly*t_plate (z in [0,t]) with n_stiff flat bars
 * web_h*web_t along x on z = t, as one conforming masked Kuhn grid. */
shipfem_status shipfem_mesh_panel(double lx, double ly, double t_plate, int32_t n_stiff,
                                  double web_h, double web_t, int32_t nx, int32_t ny_bay,
                                  int32_t nz_web, shipfem_mesh** out);
shipfem_status shipfem_mesh_counts(const shipfem_mesh* m, uint32_t* nodes, uint32_t* tets);
/* Borrowed views of the mesh's own arrays (valid until the next call on m). */
shipfem_status shipfem_mesh_arrays(const shipfem_mesh* m, const double** rest_position,
                                   const uint32_t** index, const uint32_t** node_flags,
                                   const double** load);
/* OR `flags` into every node with |X[axis] - value| <= tol. */
shipfem_status shipfem_mesh_flag_plane(shipfem_mesh* m, int32_t axis, double value, double tol,
                                       uint32_t flags);
/* OR `flags` into nodes with |X[axis]-value| <= tol within `radius` of (cx, cy)
 * measured in the two other axes (used to mark indenter candidates). */
shipfem_status shipfem_mesh_flag_disc(shipfem_mesh* m, int32_t axis, double value, double tol,
                                      double c0, double c1, double radius, uint32_t flags);
/* Consistent load of a uniform traction (N/m^2) on every tet face lying in the
 * plane X[axis] == value: traction * area / 3 per face node; marks LOADED. */
shipfem_status shipfem_mesh_traction(shipfem_mesh* m, int32_t axis, double value, double tol,
                                     const double traction[3]);
/* Fill the mesh fields of a case descriptor (pointers borrowed from m). */
shipfem_status shipfem_mesh_fill_case(const shipfem_mesh* m, shipfem_case_desc* d);
void shipfem_mesh_destroy(shipfem_mesh* m);

/* Reference stable step: safety * shortest edge / sqrt(E/rho) (fem.cpp:160-185). */
double shipfem_edge_timestep(const double* rest_position, const uint32_t* index,
                             uint32_t tet_count, double youngs_modulus, double density,
                             double safety);
/* The stable step the solver uses instead of criticalTimestep (fem.cpp:183-185), without a
 * solver: cfl = 2 / omega_max with omega_e^2 = 4 lambda_max(B^T C B) / rho the largest
 * lumped-mass frequency of tet e (Irons-Hughes), evaluated in FP64 on device 0 from the
 * rest coordinates [3N] and connectivity [4T]. Multiply by a safety factor (the solver
 * takes 0.9) for dt. Uses device memory of its own for the call. */
shipfem_status shipfem_stable_timestep(const double* rest_position, const uint32_t* index,
                                       uint32_t node_count, uint32_t tet_count,
                                       double youngs_modulus, double poisson_ratio, double density,
                                       double* cfl, char* err, size_t err_len);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * evaluation and result as shipfem_stable_timestep, bit for bit, with no allocation. It
 * stages the mesh in a 32 MiB device array of the library's own (one per GPU, held until
 * the process ends) and runs on the null stream of the current device; a mesh needing more
 * (24 bytes per node + 16 per tet + 8 per 64 tets, each part rounded up to 256) is refused
 * with SHIPFEM_INVALID_ARG and a message. Calls on one GPU share that array: calls that
 * overlap on one GPU (from several host threads) race and must be serialised by the caller. */
shipfem_status shipfem_stable_timestep_noalloc(const double* rest_position, const uint32_t* index,
                                               uint32_t node_count, uint32_t tet_count,
                                               double youngs_modulus, double poisson_ratio,
                                               double density, double* cfl, char* err, size_t err_len);

/* ---- the host FP32 step (replaces sim::fem::tetForces and stepCpu) -------------- */
/* tetForces (fem.hpp:84-86, fem.cpp:73-117): the four nodal forces [12] of tet `tet` at
 * the current positions [3N], with rest_inv [9T] column-major (fem.cpp:10-16) and
 * rest_volume [T] from computeRestState. With rest_position [3N] (the rest coordinates in
 * double; may be NULL) the deformation gradient is formed from displacements, as the
 * device pass forms it; without, from the positions. Co-rotational linear elastic, the
 * rotation by four Higham iterations. */
shipfem_status shipfem_cpu_tet_forces(const float* position, const double* rest_position,
                                      const uint32_t* index, const float* rest_inv,
                                      const float* rest_volume, uint64_t tet, float lambda, float mu,
                                      float out[12]);
/* stepCpu (fem.hpp:77-82, fem.cpp:187-215): one explicit step in place on the caller's
 * TetMesh arrays. Every tet's forces go to tet_force [12T]; then per node, fixed (may be
 * NULL) nonzero: velocity 0 and position held; otherwise f = its corner forces gathered in
 * adjacency (CSR, entry = tet * 4 + corner) order minus gravity * mass along z, and
 * v = (v + f dt / mass) damping, x += v dt. Single-threaded, deterministic. */
shipfem_status shipfem_cpu_step(uint32_t node_count, uint32_t tet_count, float* position, float* velocity,
                                const float* mass, const uint32_t* fixed, const uint32_t* index,
                                const float* rest_inv, const float* rest_volume, float* tet_force,
                                const uint32_t* adjacency_offset, const uint32_t* adjacency_entry,
                                const double* rest_position, float lambda, float mu, float dt,
                                float gravity, float damping);

/* Element-pass variant. TILED (default): 256-tet tiles stage their nodes in LDS and
 * sum corner forces per tile node before one write; SLOTS: one 48-byte force slot per
 * tet gathered through the CSR node map, the reference's own scheme
 * (tet_forces.comp + node_integrate.comp). The two sum a node's forces in a
 * different association, so their FP32 results differ in the last bits. */
#define SHIPFEM_VARIANT_TILED 0
#define SHIPFEM_VARIANT_SLOTS 1

/* Launch shape. For a given variant every choice gives bit-identical results:
 * reductions are fixed per 64-item wavefront and per 256-tet tile of each case,
 * independent of block size and order. */
typedef struct shipfem_launch_config {
    int32_t elem_block; /* SLOTS: threads per element-pass block (64, 128, 256, 512) */
    int32_t node_block; /* threads per node-pass block (64, 128, 256, 512) */
    int32_t xcd_remap;  /* 1: give each XCD a contiguous range of tiles */
    int32_t chunk;      /* steps per captured graph (even) */
    int32_t variant;    /* SHIPFEM_VARIANT_* */
    int32_t split;      /* a batch of two or more cases: 0 runs them as up to four parts
                         * (balanced by element count) on as many streams, each replaying its
                         * own captured chunks, so each part's kernels fill the others' launch
                         * gaps and tails (each case's arithmetic unchanged); -1: one stream */
    int32_t rest_from_positions; /* tiled variant: 1 forms each tet's Dm^-1 and V in the pass from
                         * its tile's rest positions (16 B per tile node read instead of 40 B
                         * per tet, ~40 more FP32 operations a tet, results within ~1e-6 of the
                         * stored planes'); 0 reads the stored planes */
    int32_t resident;   /* 0: a plan whose tet cases each fit one block (at most 1024 tets in
                         * four tiles, 2048 nodes) runs each step's passes inside one
                         * 1024-thread block per case, a whole chunk per launch (same bits);
                         * -1: two launches per step */
} shipfem_launch_config;

/* ---- one case (replaces gpu::FemGpuSolver, fem_gpu.hpp:17-45) -------------- */
typedef struct shipfem_solver shipfem_solver;

/* initialise (fem_gpu.hpp:23-24): computes the rest state (FP64 host, stored
 * FP32), uploads SoA planes, evaluates the FP64 element-frequency time step on the
 * device. */
shipfem_status shipfem_solver_create(const shipfem_case_desc* desc, shipfem_solver** out,
                                     char* err, size_t err_len);
/* cfl = 2/omega_max (FP64 device min over elements), dt in use, and the
 * reference's edge formula at safety 0.5 for comparison. */
shipfem_status shipfem_solver_timestep(const shipfem_solver* s, double* cfl, double* dt,
                                       double* edge_dt);
/* Changes dt/gravity/damping/E/nu for the following runs (the per-call arguments
 * of fem_gpu.hpp:31-32). */
shipfem_status shipfem_solver_set_params(shipfem_solver* s, const shipfem_step_params* p);
/* run (fem_gpu.hpp:31-32): advances `steps` steps and returns the device time in
 * ms measured with HIP events on the solver's stream. mode 0 replays captured
 * hipGraph chunks; mode 1 issues two plain launches per step, as the reference
 * records two dispatches per step (fem_gpu.cpp:401-413). */
shipfem_status shipfem_solver_run(shipfem_solver* s, int32_t steps, int32_t mode,
                                  double* device_ms);
/* readback (fem_gpu.hpp:34-35). */
shipfem_status shipfem_solver_readback(shipfem_solver* s, shipfem_state_view* view);
/* Energy records written so far (one per sample step). */
shipfem_status shipfem_solver_energy(shipfem_solver* s, shipfem_energy_record* out,
                                     int32_t max_records, int32_t* count);
/* Per-tet nodal forces of the last element pass, reference layout [12T]
 * (tetForce, fem.hpp:49-50). SLOTS variant only (SHIPFEM_UNSUPPORTED otherwise). */
shipfem_status shipfem_solver_tet_forces(shipfem_solver* s, float* out);
/* Launch shape and element-pass variant of this solver (zeros keep the defaults:
 * tiled, 256-thread node blocks, XCD remap on). */
shipfem_status shipfem_solver_set_launch(shipfem_solver* s, const shipfem_launch_config* cfg);
/* Runs `steps` plain steps on the solver's stream, each element and node launch
 * carrying its own start/stop events (recorded by the dispatch, so the elapsed
 * time is the kernel's execution), and returns the mean element-pass and node-pass
 * durations in ms (the per-kernel split of the timings fem_gpu.cpp:388-421 takes
 * around a whole run). */
shipfem_status shipfem_solver_profile(shipfem_solver* s, int32_t steps, double* elem_ms,
                                      double* node_ms);
/* Device bytes held by the solver. */
shipfem_status shipfem_solver_footprint(const shipfem_solver* s, uint64_t* bytes);
/* The HBM bytes one step's passes must move, by the layout of this case: out[0], out[1]
 * the element pass's reads and writes, out[2], out[3] the node pass's. Every plane a pass
 * reads is counted once (a node's displacement once, however many tiles gather it) and
 * every plane it writes once. A J2 tet pass reads every tet's eqps and the plastic strain
 * (and damage) only of tets that have yielded, counted on the device at the call. Not
 * counted are the J2 history a yielding tet writes back and the per-wave FP64 sums. The
 * roofline of a pass is these bytes over its duration. */
shipfem_status shipfem_solver_traffic(const shipfem_solver* s, uint64_t out[4]);
void shipfem_solver_destroy(shipfem_solver* s);

/* One torn tet: the step whose element pass tore it and the failure plane, the
 * maximum principal direction of its world-frame stress at the tear with the largest
 * component positive (the reference's Increment::failureNormal, plasticity.cpp:472,
 * taken in the world frame). */
typedef struct shipfem_tear {
    uint32_t tet;
    int64_t step;
    double normal[3];
} shipfem_tear;

/* The tears so far, ordered by step then tet; *count is their total (it may exceed
 * max). Empty for a case without failure. */
shipfem_status shipfem_solver_tears(shipfem_solver* s, shipfem_tear* out, uint32_t max,
                                    uint32_t* count);

/* Bytes a tet case of this shape will hold on the device, for meshes whose nodes lie in at
 * most 32 tets (Kuhn-type and panel meshes). plastic: 0 elastic, 1 J2, 2 J2 with failure. */
uint64_t shipfem_case_footprint(uint32_t node_count, uint32_t tet_count, int32_t plastic,
                                int32_t sample_capacity);
/* An upper bound for any tet mesh: the same formula with the mesh's own largest node
 * degree and up to four tile slots per tet (0 for an empty descriptor). The server refuses
 * at submit a case whose bound exceeds its budget. */
uint64_t shipfem_case_footprint_mesh(const shipfem_case_desc* desc, int32_t sample_capacity);

/* The mesh-regularised failure strain (plasticity.cpp:158-169): an element of in-plane
 * size `length` in a plate of `thickness` reads the neck diluted by t/l,
 * uniform + (fracture - uniform) * min(thickness / length, 1). */
double shipfem_regularised_failure_strain(double uniform, double fracture, double length,
                                          double thickness);

/* The groups (panels) whose torn volume reaches `fraction` of the group's volume
 * (Solver::tornPanelsAt, zone.cpp:1262-1275, by volume for tets), ascending by id.
 * tet_group[T] labels each tet; damage[T] is a readback (1 = torn). *count is the
 * number of such groups (it may exceed max). */
shipfem_status shipfem_torn_groups(const double* rest_position, const uint32_t* index,
                                   uint32_t tet_count, const uint32_t* tet_group,
                                   const float* damage, double fraction, uint32_t* out,
                                   uint32_t max, uint32_t* count);

/* ---- hand-off to and from other solvers, checkpoint, driven edge ----------- */

/* Everything the device holds, back in double, in the shape zone::Solver::adopt takes
 * (gpu::ZoneGpuState, engine/gpu/zone_gpu.hpp:113-126; Solver::adopt,
 * engine/sim/zone.hpp:963-969). The caller owns the arrays; any may be NULL.
 * The account's running sums are those of the steps taken so far: an energy record at
 * step n adds the half of step n's work that lies before t_n, so a record and these
 * sums agree only at the steps a record is taken after. */
typedef struct shipfem_handoff {
    double* position;     /* [3N] X + u */
    double* displacement; /* [3N] u itself (adopt prefers it to position - X: exact) */
    double* velocity;     /* [3N] v_{n-1/2} */
    double* plastic;      /* [8T] per tet: eps_p (6, Voigt, engineering shears, co-rotated
                           * frame), eqps, damage (1 = torn) */
    int64_t* tear_step;   /* [T] step whose element pass tore the tet, -1 intact */
    double* tear_stress;  /* [6T] world-frame stress at the tear, Voigt */
    int64_t steps;        /* n: the state holds u_n and v_{n-1/2} */
    double time;          /* n dt */
    double penetration;   /* indenter travel |v_ind| t; 0 without an indenter */
    double work;          /* w_ext + w_contact + w_boundary */
    double w_ext, w_contact, w_boundary, damping_loss, dissipation, fracture;
    uint32_t torn;        /* tets torn so far */
    /* A plastic solid-shell hex keeps its history per integration point
     * (solidshell::ElementPlasticState, engine/sim/solid_shell.hpp:436-441): */
    double* hex_points;   /* [64T] per element, per point g = 0..7 (solid_shell.cpp's Gauss
                           * order), 8 values: eps_p (6, Voigt, engineering shears,
                           * co-rotated frame), eqps, damage (1 = the point has failed; an
                           * element whose eight points have failed is torn). The back stress
                           * is zero: hardening is isotropic (kinematicModulus 0). */
    double* hex_enhanced; /* [7T] the enhanced-strain parameters alpha the next update
                           * starts its Newton iterations from */
} shipfem_handoff;

/* ZoneGpuSolver::readback (engine/gpu/zone_gpu.hpp:205-207): the state and the
 * account's running sums, for a CPU solver to adopt and continue. */
shipfem_status shipfem_solver_export(shipfem_solver* s, shipfem_handoff* out);

/* zone::Solver::adopt (engine/sim/zone.hpp:963-969, zone.cpp:1321-1368) on this
 * solver: take a state computed elsewhere and continue as if this solver had produced
 * it. velocity and position (or displacement) are required; plastic (and with it tear_step/tear_stress)
 * is taken when non-NULL and the case is a plastic tet case; for a plastic hex case
 * hex_points (and with it hex_enhanced, zeros when NULL, and tear_step) takes its place,
 * and a plastic hex exports its per-element summary in plastic (eps_p zero, the largest
 * eqps and the smallest damage of its points: readback's eqps and damage). steps sets the step counter and time
 * must equal steps * dt (to 1e-9 relative), else SHIPFEM_INVALID_ARG. displacement,
 * when non-NULL, is taken instead of position - X (a double u of an FP32 + compensation
 * state comes back to the same two floats, so a run adopting its own export continues
 * with the same bits; position - X loses u's last bits to X's magnitude). The running
 * sums continue from w_ext .. fracture. Nothing is stepped, so the plastic history
 * comes back exactly as handed over. The energy history restarts empty. */
shipfem_status shipfem_solver_adopt(shipfem_solver* s, const shipfem_handoff* in);

/* Bitwise checkpoint of the solver's evolving device state (displacement and velocity
 * with their compensations, plastic and tear state, the account's per-wave sums, step
 * counter, status and energy history). With buf NULL, *len receives the size. A solver
 * of the same case restored from it continues with the same bits as the run that wrote
 * it. No reference counterpart: the reference has no checkpoint of a GPU solve. */
shipfem_status shipfem_solver_checkpoint(shipfem_solver* s, void* buf, uint64_t* len);
/* SHIPFEM_INVALID_ARG when buf was written by a solver of another case shape,
 * material features or time step. */
shipfem_status shipfem_solver_restore(shipfem_solver* s, const void* buf, uint64_t len);

/* coupling::applyEdgeDrive (engine/sim/coupling.hpp:256-270, coupling.cpp:110-125):
 * drives every DOF whose driven[3 n + a] is non-zero to displacement[3 n + a] (Tier-1's
 * interface displacement, on the drive_ramp smoothstep) by setting its DRIVE bit,
 * clearing its PIN bit (the drive replaces a zero clamp) and writing prescribed. It
 * never un-pins: a DOF the flags pin and the drive does not name stays clamped.
 * Returns the number of DOF written. */
uint64_t shipfem_apply_edge_drive(uint32_t node_count, const double* displacement,
                                  const uint8_t* driven, uint32_t* node_flags,
                                  double* prescribed);

/* ---- case wire format ------------------------------------------------------- */
/* One self-checking little-endian blob per case (tets, or solid-shell hexes through the
 * shipfem_hex_case_* calls, marked in the header): the on-disk
 * case file and the form a case travels in to the server. A 64-byte header (magic "SHIPFEMC", version, node and
 * tet counts, which optional arrays follow, payload size, FNV-1a 64 of the payload),
 * the scalar block, then rest positions, connectivity and the optional arrays of the
 * descriptor. No reference counterpart: the reference builds its cases in memory; a
 * run's tables are reproducible from files it can re-read (tools/zone_gpu_probe/
 * sweep.sh:21-25 is the reason). */
typedef struct shipfem_case_blob shipfem_case_blob; /* owns a decoded case's arrays */
uint64_t shipfem_case_encoded_size(const shipfem_case_desc* d);
shipfem_status shipfem_case_encode(const shipfem_case_desc* d, void* buf, uint64_t len);
/* Fills *d with pointers into *out (valid until shipfem_case_blob_destroy).
 * SHIPFEM_INVALID_ARG on a wrong magic or version, inconsistent sizes, a checksum
 * mismatch or a node id out of range. */
shipfem_status shipfem_case_decode(const void* buf, uint64_t len, shipfem_case_blob** out,
                                   shipfem_case_desc* d);
void shipfem_case_blob_destroy(shipfem_case_blob* b);
/* The blob to and from a file (written to path.tmp, then renamed). */
shipfem_status shipfem_case_save(const shipfem_case_desc* d, const char* path);
shipfem_status shipfem_case_load(const char* path, shipfem_case_blob** out,
                                 shipfem_case_desc* d);
/* The same for a solid-shell hex case (8 node ids per element); decode and load read
 * either kind, and shipfem_case_blob_corners tells which (4 or 8). */
uint64_t shipfem_hex_case_encoded_size(const shipfem_case_desc* d);
shipfem_status shipfem_hex_case_encode(const shipfem_case_desc* d, void* buf, uint64_t len);
shipfem_status shipfem_hex_case_save(const shipfem_case_desc* d, const char* path);
uint32_t shipfem_case_blob_corners(const shipfem_case_blob* b);

/* ---- the solid-shell hexahedron (engine/sim/solid_shell.hpp:111-194) --------- */
/* An 8-node hex whose nodes 0-3 are the zeta = -1 face and 4-7 the zeta = +1 face,
 * each counter-clockwise seen from +zeta (solid_shell.hpp:115-125): zeta is the
 * shell's thickness. Assumed natural transverse shear and thickness strain, seven
 * enhanced strains condensed at rest, 2x2x2 Gauss. Host FP64. `nodes` is 8 x xyz. */

/* Condensed elastic stiffness, 24 x 24 row-major, global frame
 * (solidshell::elementStiffness with Formulation::SolidShell, solid_shell.hpp:202-203).
 * SHIPFEM_INVALID_ARG for a degenerate or inverted element. */
shipfem_status shipfem_hex_stiffness(const double nodes[24], double youngs_modulus,
                                     double poisson_ratio, double out[576]);
/* The failure strain initialisePlasticState resolves for one element (solid_shell.cpp:808-815,
 * elementSize :770-806, plasticity::regularisedFailureStrain plasticity.cpp:158-169): in-plane
 * size l = sqrt(mid-surface area), thickness t = volume / area, and
 * uniform + (fracture - uniform) min(1, t / l); fracture itself for a degenerate element. */
shipfem_status shipfem_hex_failure_strain(const double nodes[24], double uniform, double fracture,
                                          double* out);
/* Lumped mass rho int N_a dV per node (solidshell::elementMass, solid_shell.hpp:330). */
shipfem_status shipfem_hex_mass(const double nodes[24], double density, double out[8]);
/* The rotation the element carries: the polar factor of the deformation gradient at the
 * element centre, column-major (solidshell::elementRotation, solid_shell.hpp:143). A rest
 * element with a singular centre Jacobian gives the identity and SHIPFEM_INVALID_ARG. */
shipfem_status shipfem_hex_rotation(const double rest[24], const double current[24],
                                    double out[9]);
/* One element's elastoplastic update (solidshell::elementPlasticUpdate with
 * Formulation::SolidShell, solid_shell.hpp:311-316), on the host, by the same element code
 * the plastic hex solve runs on the device: co-rotated at the centre's polar factor, the
 * seven enhanced parameters solved by Newton on the algorithmic tangent, J2 with linear
 * hardening at each Gauss point, Rice-Tracey damage when failure_strain > 0, the step re-run
 * without enhanced modes when a point fails. point_state[8 * 8]: per point the plastic
 * strain (6, engineering shears), eqps and damage (1 = failed); enhanced[7]: the warm
 * start. Both are advanced in place. force: what the element applies to its nodes. An
 * element whose eight points have all failed is left as it is, with a zero force. */
shipfem_status shipfem_hex_plastic_update(const double rest[24], const double current[24],
                                          double youngs_modulus, double poisson_ratio,
                                          double yield_strength, double hardening,
                                          double failure_strain, double sensitivity,
                                          double reference_triaxiality, double cutoff_triaxiality,
                                          double point_state[64], double enhanced[7],
                                          double force[24], double* dissipation,
                                          int32_t* failed_points);
/* Co-rotational internal force on the nodes, f = -R K (R^T x - X), R the polar factor
 * of the deformation gradient at the element centre (solidshell::internalForce,
 * solid_shell.hpp:355-356). */
shipfem_status shipfem_hex_force(const double stiffness[576], const double rest[24],
                                 const double current[24], double out[24]);
/* safety * 2 / omega_max of the element with its lumped mass, omega_max by power
 * iteration (solidshell::criticalTimestep, solid_shell.hpp:369-370). */
shipfem_status shipfem_hex_critical_timestep(const double nodes[24], double youngs_modulus,
                                             double poisson_ratio, double density,
                                             double safety, double* dt);

/* solidshell::makePlateMesh (solid_shell.hpp:546-552): a plate lx x ly x thickness, nx x ny
 * elements in plane and nz through it, nodes numbered with the thickness index fastest,
 * then y, then x, mid-surface at z = 0. Counts only when rest_position and index are
 * NULL; otherwise rest_position holds 3 * node_count and index 8 * element_count. */
shipfem_status shipfem_hex_plate_mesh(double lx, double ly, double thickness, int32_t nx,
                                      int32_t ny, int32_t nz, uint32_t* node_count,
                                      uint32_t* element_count, double* rest_position,
                                      uint32_t* index);
/* solidshell::uniformPressureLoad (solid_shell.hpp:714-719): the consistent nodal load of
 * a pressure on every element's +zeta face that lies on the mesh boundary, added into
 * load[3 * node_count]; positive pressure pushes into the element. */
shipfem_status shipfem_hex_pressure_load(const double* rest_position, uint32_t node_count,
                                         const uint32_t* index, uint32_t element_count,
                                         double pressure, double* load);

/* An explicit solve over solid-shell hexes: the zone::Solver step
 * (engine/sim/zone.cpp:928-1070, 1086-1252) on the device. The case is a
 * shipfem_case_desc read with tet_count as the element count and index as
 * [8 * tet_count] node ids in the order above.
 *   material.plastic 0: co-rotational elastic, the condensed stiffness formed at rest
 *     (solidshell::internalForce, solid_shell.cpp:735-766).
 *   material.plastic 1: elementPlasticUpdate (solid_shell.cpp:826-1029) in FP64 at eight
 *     points per element: J2 with linear hardening, the enhanced-strain Newton, ductile
 *     damage when failure_strain > 0 (taken as given for every element; the reference
 *     regularises it per element size, see shipfem_regularised_failure_strain), and torn
 *     elements (every point failed) carrying nothing.
 * dt = 0 takes dt_safety * 2 / omega_max, omega_max the largest element frequency (FP64,
 * host). The returned handle takes every shipfem_solver_* call;
 * shipfem_solver_tet_forces returns [24 * tet_count] (8 corners x xyz per element). The
 * readback's per-element eqps is the largest of the element's points, plastic_strain
 * their mean and damage the smallest (1: torn); tears carry no plane (zeros).
 * shipfem_solver_export / adopt refuse a plastic hex case (SHIPFEM_UNSUPPORTED: eight
 * points of history per element); checkpoint and restore carry it. */
shipfem_status shipfem_hex_solver_create(const shipfem_case_desc* desc, shipfem_solver** out,
                                         char* err, size_t err_len);

/* ---- many cases in one launch per phase ----------------------------------- */
typedef struct shipfem_batch shipfem_batch;


shipfem_status shipfem_batch_create(const shipfem_case_desc* cases, uint32_t case_count,
                                    const shipfem_launch_config* cfg, shipfem_batch** out,
                                    char* err, size_t err_len);
/* shipfem_batch_create for solid-shell hex cases (each read as shipfem_hex_solver_create
 * reads its case); every batch call then applies. A batch is all elastic or all plastic. */
shipfem_status shipfem_hex_batch_create(const shipfem_case_desc* cases, uint32_t case_count,
                                        const shipfem_launch_config* cfg, shipfem_batch** out,
                                        char* err, size_t err_len);
/* All cases advance `steps` (cases that reach max_steps stop). */
shipfem_status shipfem_batch_run(shipfem_batch* b, int32_t steps, double* device_ms);
/* Times candidate launch shapes on the batch itself (state is snapshotted and
 * restored) and keeps the fastest; writes it to *best. */
shipfem_status shipfem_batch_tune(shipfem_batch* b, int32_t steps_per_trial,
                                  shipfem_launch_config* best, double* best_ms_per_step);
shipfem_status shipfem_batch_set_config(shipfem_batch* b, const shipfem_launch_config* cfg);
/* shipfem_solver_profile for a batch: `steps` (whole chunks, from a chunk boundary) plain
 * steps with per-dispatch events; mean element-pass and node-pass kernel durations. */
shipfem_status shipfem_batch_profile(shipfem_batch* b, int32_t steps, double* elem_ms, double* node_ms);
/* Device bytes of the batch's case arenas (one allocation; creation copies it once,
 * device to device, as the state reset() restores). */
shipfem_status shipfem_batch_footprint(const shipfem_batch* b, uint64_t* bytes);
/* shipfem_solver_traffic summed over the batch's cases. */
shipfem_status shipfem_batch_traffic(const shipfem_batch* b, uint64_t out[4]);
shipfem_status shipfem_batch_readback(shipfem_batch* b, uint32_t case_index,
                                      shipfem_state_view* view);
shipfem_status shipfem_batch_energy(shipfem_batch* b, uint32_t case_index,
                                    shipfem_energy_record* out, int32_t max_records,
                                    int32_t* count);
shipfem_status shipfem_batch_tears(shipfem_batch* b, uint32_t case_index, shipfem_tear* out,
                                   uint32_t max, uint32_t* count);
/* Restart every case from its initial state (for repeated timing). */
shipfem_status shipfem_batch_reset(shipfem_batch* b);
void shipfem_batch_destroy(shipfem_batch* b);

/* ---- resident serving loop with continuous batching ---------------------- */
typedef struct shipfem_server shipfem_server;

/* Which resident cases a round advances. ALL: every one, so each case's latency is its
 * round count times rounds sized by everything resident. SHORTEST: the cases with the least
 * remaining work (tets x steps left) first, until the round holds round_work tet-steps; a
 * case left out of starve_rounds consecutive rounds of its kind goes first, so every case
 * advances. The policy changes only when a case runs, never its arithmetic. */
#define SHIPFEM_ROUNDS_ALL 0
#define SHIPFEM_ROUNDS_SHORTEST 1

typedef struct shipfem_server_opts {
    uint64_t memory_budget;  /* device bytes for case state; 0 => 60% of HBM */
    uint32_t max_active;     /* slots; 0 => 256 */
    int32_t chunk;           /* steps per device round between admissions; 0 => 64 */
    shipfem_launch_config launch; /* zeros => defaults */
    int32_t policy;          /* SHIPFEM_ROUNDS_* ; 0 => ALL */
    double round_work;       /* SHORTEST: tet-steps a round takes (at least one case); 0 => 2e8 */
    int32_t starve_rounds;   /* SHORTEST: 0 => 4 */
} shipfem_server_opts;

typedef struct shipfem_result {
    uint64_t request_id;
    shipfem_status status;
    int64_t steps;
    double sim_time;
    double submit_s, admit_s, done_s; /* host monotonic clock, seconds */
    uint32_t node_count, tet_count;
    double max_displacement;  /* max |u| over nodes, m */
    double max_eqps;
    double peak_contact_force; /* max over records of |sum contact force|, N */
    uint32_t torn_elements;    /* tets torn by the end (failure cases) */
    double torn_volume;        /* their rest volume, m^3 */
    shipfem_energy_record last; /* energy at the last record */
} shipfem_result;

typedef struct shipfem_server_stats {
    uint64_t submitted, admitted, completed, failed, refused;
    uint64_t active, queued;
    uint64_t bytes_in_use, bytes_budget;
    uint64_t rounds;
    double element_steps; /* total tet-steps advanced */
    double device_ms;     /* device time of all rounds */
} shipfem_server_stats;

shipfem_status shipfem_server_create(const shipfem_server_opts* opts, shipfem_server** out,
                                     char* err, size_t err_len);
/* Copies the case; runs it to `run_to_time` seconds of simulated time (rounded up
 * to whole steps). Refuses (SHIPFEM_BUSY) a case whose footprint exceeds the budget. */
shipfem_status shipfem_server_submit(shipfem_server* s, const shipfem_case_desc* desc,
                                     double run_to_time, uint64_t* request_id);
/* shipfem_server_submit for a solid-shell hex case, read as shipfem_hex_solver_create
 * reads it (elastic or plastic). Rounds run one element kind at a time (tets, elastic
 * hexes, plastic hexes, in turn). The footprint is measured after the case's set-up: one
 * that can never fit the budget completes with SHIPFEM_BUSY in its result. */
shipfem_status shipfem_server_submit_hex(shipfem_server* s, const shipfem_case_desc* desc,
                                         double run_to_time, uint64_t* request_id);
/* shipfem_server_submit of an encoded case (the wire format above). */
shipfem_status shipfem_server_submit_encoded(shipfem_server* s, const void* buf, uint64_t len,
                                             double run_to, uint64_t* request_id);
/* Moves up to `max` completed results into out. Optionally keeps the final
 * displacement of request ids the caller asked for with shipfem_server_keep. */
shipfem_status shipfem_server_poll(shipfem_server* s, shipfem_result* out, uint32_t max,
                                   uint32_t* count);
/* Ask that the final state of request_id be retained for shipfem_server_fetch. */
shipfem_status shipfem_server_keep(shipfem_server* s, uint64_t request_id);
/* The kept final state of a completed request, once: every field of the view that is
 * not NULL (displacement, velocity, position X + u, eqps, plastic strain, damage), in the
 * caller's node and element orders, as shipfem_solver_readback returns it.
 * SHIPFEM_NOT_FOUND for a request not kept, not finished or already fetched. */
shipfem_status shipfem_server_fetch(shipfem_server* s, uint64_t request_id,
                                    shipfem_state_view* view);
/* Blocks until every submitted request has completed or `timeout_s` passes. */
shipfem_status shipfem_server_drain(shipfem_server* s, double timeout_s);
shipfem_status shipfem_server_stats_get(shipfem_server* s, shipfem_server_stats* out);
double shipfem_now(void); /* the server's monotonic clock, seconds */
void shipfem_server_destroy(shipfem_server* s);

#ifdef __cplusplus
}
#endif
#endif /* SHIPFEM_H */
