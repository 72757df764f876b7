/*
 * oracle.h -- CPU oracle for the explicit tetrahedral FEM step.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
 * code: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker (or, for the FP32 reference restatement, as the
 * timed CPU baseline). The product path fails loudly without its HIP library.
 *
 * Two restatements live here:
 *
 *   1. The FP64 oracle (ora_*): the reference's co-rotational tet step
 *      (engine/sim/fem.cpp:73-117 tetForces, :187-215 stepCpu) in double, extended
 *      with the pieces the reference specifies on other paths:
 *        - J2 radial return, linear isotropic hardening, one point per tet
 *          (engine/sim/plasticity.cpp:319-455);
 *        - ramped dead loads (engine/sim/zone.hpp:579-589) and per-DOF pins and a
 *          smoothstep prescribed drive (engine/sim/zone.cpp:1080-1145);
 *        - a rigid spherical penalty indenter (new; the reference has only a flat
 *          gripping punch, engine/sim/zone.cpp:1147-1198);
 *        - the energy account modelled on engine/sim/zone.cpp:1014-1070.
 *   2. The FP32 reference restatement (ref32_*): the exact float operation order
 *      of engine/sim/fem.cpp, bit-identical to the compiled reference when both are
 *      built with -ffp-contract=off. Single thread and a static all-core split.
 *
 * Parity status of every function is in the header comment of oracle.c.
 */
#ifndef SHIPFEM_ORACLE_H
#define SHIPFEM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- node flags (shared meaning with the product's include/shipfem.h) ---- */
enum {
    ORA_PIN_X = 1u << 0,
    ORA_PIN_Y = 1u << 1,
    ORA_PIN_Z = 1u << 2,
    ORA_LOADED = 1u << 3,
    ORA_CONTACT = 1u << 4,
    ORA_DRIVE_X = 1u << 5,
    ORA_DRIVE_Y = 1u << 6,
    ORA_DRIVE_Z = 1u << 7,
};

/* ---- meshes --------------------------------------------------------------- */

/* Kuhn box, reference node numbering and tet order (engine/sim/fem.cpp:217-263),
 * in double. Returns counts through nodes/tets when X/index are NULL. */
void ora_box_mesh(double lx, double ly, double lz, int nx, int ny, int nz, int64_t* nodes,
                  int64_t* tets, double* X, uint32_t* index);

/* Same mesh built exactly as the reference does it, in float (positions
 * lx*float(i)/float(nx), float orientation determinant). */
void ora_box_mesh_f32(float lx, float ly, float lz, int nx, int ny, int nz, float* X,
                      uint32_t* index);

/* Counter-based uniform in [0,1): SplitMix64 finaliser of seed + (counter+1)*phi. */
double ora_uniform(uint64_t seed, uint64_t counter);

/* Jitter lattice node coordinates of a box mesh by amp*h_a*U(-1,1) per axis,
 * except the component normal to a boundary face the node lies on. */
void ora_box_jitter(double* X, int nx, int ny, int nz, double lx, double ly, double lz,
                    double amp, uint64_t seed);

typedef struct ora_panel_spec {
    double lx, ly, t_plate; /* plate: x in [0,lx], y in [0,ly], z in [0,t_plate] */
    int n_stiff;            /* flat bars along x, centred at y = ly*(s+1)/(n_stiff+1) */
    double web_h, web_t;    /* bar height (z) and thickness (y) */
    int nx, ny_bay, nz_web; /* cells along x, per bay in y, and up the web */
} ora_panel_spec;

/* Stiffened panel as a masked Kuhn grid: every cell of the plate layer, plus the
 * web columns above it. Counts only when X/index are NULL. */
void ora_panel_mesh(const ora_panel_spec* s, int64_t* nodes, int64_t* tets, double* X,
                    uint32_t* index);

/* ---- rest state (engine/sim/fem.cpp:119-158, in double) ------------------- */
void ora_rest_state(const double* X, int64_t N, const uint32_t* index, int64_t T, double density,
                    double* rest_inv /*9T col-major*/, double* volume /*T*/, double* mass /*N*/,
                    uint32_t* adj_offset /*N+1*/, uint32_t* adj_entry /*4T*/);

/* Reference stable step: safety * shortest edge / sqrt(E/rho) (fem.cpp:160-185). */
double ora_edge_timestep(const double* X, const uint32_t* index, int64_t T, double E, double rho,
                         double safety);

/* Largest element frequency of the lumped-mass linear tet: omega_e^2 =
 * lambda_max(K_e)/(rho V/4), K_e = V B^T C B. Returns max_e omega_e; the stable
 * central-difference step is 2/omega (Irons-Hughes element bound). */
double ora_max_element_omega(const double* rest_inv, int64_t T, double E, double nu, double rho);

/* ---- loads ------------------------------------------------------------------ */
/* Consistent nodal loads of a uniform traction t (N/m^2) over every tet face whose
 * three nodes all lie on the plane X[axis] == value (|d| <= tol): t*A/3 per node. */
void ora_face_traction(const double* X, const uint32_t* index, int64_t T, int axis, double value,
                       double tol, const double traction[3], double* load /*3N, accumulated*/);

/* ---- FP64 element and point law ----------------------------------------- */
void ora_tet_force(const double x[12] /*4 nodes xyz*/, const double rest_inv[9], double volume,
                   double lambda, double mu, double out[12]);

/* Co-rotational force of one solid-shell hex, f = -R K (R^T x - X) with R the polar
 * factor of the centre deformation gradient (solid_shell.cpp:680-766). */
void ora_hex_force(const double K[576], const double X[24], const double x[24], double out[24]);

typedef struct ora_j2 {
    double E, nu, yield, hardening; /* linear isotropic hardening */
    /* ductile failure (plasticity.hpp:160-210); failure_strain <= 0 disables it */
    double failure_strain;          /* eps_f, already regularised for the element */
    double sensitivity;             /* Rice-Tracey: eps_f exp(-sensitivity (eta - eta_ref)) */
    double reference_triaxiality;   /* eta_ref (1/3, uniaxial tension) */
    double cutoff_triaxiality;      /* at or below: no damage */
} ora_j2;

/* elementPlasticUpdate (solid_shell.cpp:826-1029) on precomputed forms: B [8][6][24] and
 * G [8][6][7] of each Gauss point, w [8] its volume. point [8][9] per Gauss point: plastic
 * strain 6 (engineering shears), eqps, damage, failed; enhanced [7] the warm start in, the
 * solution out. m: E, nu, yield, hardening and the failure strain (<= 0 never fails) with
 * the Rice-Tracey constants. force [24] on the nodes, stress [48] (8 x 6, may be NULL),
 * info [6] = dissipation, iterations, converged, yielded, failed points, torn. Returns
 * torn. */
int ora_hex_plastic_update(const double* B, const double* G, const double* w,
                           const double X[24], const double x[24], const ora_j2* m,
                           double* point, double enhanced[7], double force[24],
                           double stress[48], double info[6]);

/* Radial return (plasticity.cpp:319-455, kinematic 0, no damage). strain and
 * eps_p are Voigt [xx,yy,zz,xy,yz,zx] with engineering shears; stress has tensor
 * components. Returns 1 when the point yielded. dissipation = sigma : d eps_p. */
int ora_j2_update(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                  double stress[6], double* dissipation);

/* The same return with ductile damage (plasticity.cpp:457-477): damage grows by
 * d eqps / (eps_f * triaxiality factor) and the point fails at damage 1, after which it
 * carries no stress and its history is frozen. damage >= 1 is the failed state.
 * Returns bit 0 yielded, bit 1 failed in this call. At failure, normal (may be NULL)
 * receives the maximum principal direction of the returned stress, largest component
 * positive. eta (may be NULL) receives the triaxiality of the returned stress. */
int ora_j2_update_damage(const ora_j2* m, const double strain[6], double eps_p[6],
                         double* eqps, double* damage, double stress[6], double* dissipation,
                         double normal[3], double* eta);

/* Maximum principal direction of a symmetric 3x3 stress (column-major), largest
 * component positive: the failure-plane normal of a torn tet in the world frame. */
void ora_world_stress_normal(const double s9[9], double normal[3]);

/* eps_u + (eps_frac - eps_u) min(t / l, 1) (plasticity.cpp:158-169). */
double ora_regularised_failure_strain(double uniform, double fracture, double length,
                                      double thickness);
/* exp(-sensitivity (eta - eta_ref)), infinite at or below the cutoff
 * (plasticity.cpp:171-174). */
double ora_triaxiality_factor(double sensitivity, double reference, double cutoff, double eta);

/* ---- FP64 explicit solver --------------------------------------------------- */
typedef struct ora_case {
    int64_t N, T;
    const double* X;          /* 3N rest positions */
    const uint32_t* index;    /* 4T */
    const uint32_t* flags;    /* N node flags (ORA_*), may be NULL */
    const double* load;       /* 3N full-ramp nodal load, may be NULL */
    const double* prescribed; /* 3N prescribed displacement of driven DOF, may be NULL */
    double E, nu, rho;
    int plastic;
    double yield, hardening;
    double load_ramp;  /* s(t) = min(t/load_ramp, 1); 0 => 1 */
    double drive_ramp; /* smoothstep of t/drive_ramp; 0 => 1 */
    double gravity, damping;
    /* ductile failure when plastic and failure_strain > 0 (see ora_j2) */
    double failure_strain, sensitivity, reference_triaxiality, cutoff_triaxiality;
    int indenter;
    double ind_radius, ind_center[3], ind_velocity[3], ind_kappa;
    double dt;
    /* derived (filled by ora_case_init) */
    double* rest_inv;
    double* volume;
    double* mass;
    uint32_t* adj_offset;
    uint32_t* adj_entry;
    /* Solid-shell hex case when corners == 8 (0 or 4: tets): index holds 8 node ids per
     * element (T elements), hex_k [T][576] the condensed element stiffness and hex_mass
     * [T][8] the element masses, both from the caller; elastic only. */
    int corners;
    const double* hex_k;
    const double* hex_mass;
    /* plastic hexes: the forms of every element (B [T][8][6][24], G [T][8][6][7], Gauss
     * volumes w [T][8]) */
    const double* hex_B;
    const double* hex_G;
    const double* hex_w;
    /* plastic hexes: [T] each element's failure strain (initialisePlasticState); NULL:
     * every element takes failure_strain */
    const double* hex_failure;
} ora_case;

typedef struct ora_state {
    double* u;      /* 3N */
    double* v;      /* 3N, half-step velocity */
    double* eps_p;  /* 6T */
    double* eqps;   /* T */
    double* damage; /* T, may be NULL when failure is off; >= 1 means torn */
    int64_t* tear_step;   /* T, step at which the tet tore, -1 intact (with damage) */
    double* tear_normal;  /* 3T, world-frame failure-plane normal (with damage) */
    double* force;  /* 12T (hex: 24T) scratch */
    int64_t step;   /* n: the state holds u_n, v_{n-1/2} */
    /* energy account: running sums over the steps taken so far */
    double w_ext, w_contact, w_boundary, dissipation, damping_loss;
    double fracture; /* stored energy released by torn tets, 1/2 V sigma:(eps - eps_p) */
    /* plastic hexes: per element 8 points x (plastic strain 6, eqps, damage, failed) and
     * the 7 enhanced parameters; tear_step (above) gets the step an element tore */
    double* hex_point;    /* [T][72] */
    double* hex_enhanced; /* [T][7] */
} ora_state;

typedef struct ora_energy {
    int64_t step;
    double time;
    double strain, kinetic, dissipation, w_ext, w_contact, w_boundary, damping_loss;
    double fracture;
    double contact_force[3]; /* sum of contact forces on the structure */
} ora_energy;

/* Allocates and fills the derived arrays (rest state). Returns 0 on success. */
int ora_case_init(ora_case* c);
void ora_case_free(ora_case* c);

/* Advances `steps` steps. When `sample_every` > 0, a record is written for every
 * step n with n % sample_every == 0 (energies of state n) into rec[*nrec++] until
 * max_rec. */
void ora_run(const ora_case* c, ora_state* s, int64_t steps, int64_t sample_every,
             ora_energy* rec, int64_t max_rec, int64_t* nrec);

/* ora_run on `threads` threads (static split of the element and node loops, per-thread
 * sums folded in thread order): the same per-element and per-node arithmetic, sums in
 * another order. */
void ora_run_mt(const ora_case* c, ora_state* s, int64_t steps, int64_t sample_every,
                ora_energy* rec, int64_t max_rec, int64_t* nrec, int threads);

/* ---- FP32 reference restatement (engine/sim/fem.cpp, float op order) ------- */
void ref32_rest_state(float* position, int64_t N, const uint32_t* index, int64_t T, float density,
                      float* rest_inv, float* volume, float* mass, uint32_t* adj_offset,
                      uint32_t* adj_entry);
void ref32_tet_forces(const float* position, const uint32_t* index, const float* rest_inv,
                      const float* volume, int64_t t, float lambda, float mu, float out[12]);
void ref32_step(float* position, float* velocity, const float* mass, const uint32_t* fixed,
                int64_t N, const uint32_t* index, const float* rest_inv, const float* volume,
                int64_t T, float* tet_force, const uint32_t* adj_offset,
                const uint32_t* adj_entry, float E, float nu, float dt, float gravity,
                float damping);
/* Same step with the element and node loops statically split over `threads`
 * pthreads. Bit-identical to ref32_step for any thread count. */
void ref32_step_mt(float* position, float* velocity, const float* mass, const uint32_t* fixed,
                   int64_t N, const uint32_t* index, const float* rest_inv, const float* volume,
                   int64_t T, float* tet_force, const uint32_t* adj_offset,
                   const uint32_t* adj_entry, float E, float nu, float dt, float gravity,
                   float damping, int threads);

#ifdef __cplusplus
}
#endif
#endif
