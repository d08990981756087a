/*
 * symx_scenes.h - C ABI of the example problems built on the C++ front end
 * (the counterpart of SymX's examples/NeoHookean.cpp, examples/rubber_extrusion.cpp
 * and examples/dynamic_elasticity_with_contact.cpp). Used by bench.py and the
 * parity tests: each scene defines its potentials through
 * GlobalPotential::add_potential and solves with NewtonsMethod, so every call
 * here exercises the same path a SymX user would.
 */
#ifndef SYMX_SCENES_H
#define SYMX_SCENES_H

#include <stdint.h>

#include "symx_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct symx_scene_s* symx_scene;

/* Tet4 elastic block with backward-Euler inertia and a penalty Dirichlet face. */
typedef struct symx_tet4_block_params {
    int32_t cells[3];        /* hex cells per axis (6 tets each) */
    double size[3];          /* extent; the block spans [0, size] */
    double youngs_modulus;
    double poissons_ratio;
    double density;
    double gravity[3];
    double dt;               /* <= 0: quasistatic (no inertia potential) */
    double noise;            /* uniform perturbation amplitude of free vertices */
    uint64_t seed;
    double bc_stiffness;     /* penalty 0.5 k |x - X|^2 on the fixed face */
    int32_t fixed_axis;      /* 0,1,2: vertices with coordinate == 0 on this axis are fixed; -1: none */
    int32_t material;        /* 0: Neo-Hookean, 1: stable Neo-Hookean */
    int32_t log_mu_field;    /* 1: per-tet mu = exp(log_mu[e]) parameters (see symx_scene_set_log_mu) */
    uint64_t renumber_seed;  /* != 0: vertex ids permuted at random (Fisher-Yates, SplitMix64 with this
                                seed) after the noise is applied: an unstructured numbering */
} symx_tet4_block_params;

/* Ball dropped on the ground plane z = 0 (contact scene): Tet4 jittered-grid
 * tetrahedralisation, backward-Euler inertia, IPC log-barrier per surface vertex and
 * the plane step filter in the line search. */
typedef struct symx_sphere_params {
    double radius;
    int32_t cells;             /* grid cells across the bounding cube (50: about 390k tets) */
    double jitter;             /* vertex jitter, fraction of the cell width (< 0.25) */
    uint64_t seed;
    double center[3];
    double velocity[3];        /* initial velocity of every vertex */
    double youngs_modulus;
    double poissons_ratio;
    double density;
    double gravity[3];
    double dt;
    int32_t material;          /* 0: Neo-Hookean, 1: stable Neo-Hookean */
    double contact_stiffness;  /* barrier kappa (0: no contact) */
    double dhat;               /* barrier activation distance */
    double step_filter;        /* eta of the plane step filter (0: off) */
} symx_sphere_params;

/* Hexahedral block (Hex8 or Hex27 cuboid over [0, size]; with nodes_per_element = 10 the
 * Tet10 block of the same cuboid) with a per-element Young's modulus drawn log-uniformly in
 * [youngs_min, youngs_max]; loads and boundary as the Tet4 block; element energy written
 * with fem_integrator / fem_jacobian. */
typedef struct symx_hex_block_params {
    int32_t cells[3];
    double size[3];
    int32_t nodes_per_element;  /* 8 (Hex8), 27 (Hex27) or 10 (Tet10: six per cell) */
    double youngs_min;
    double youngs_max;
    double poissons_ratio;
    double density;
    double gravity[3];
    double dt;                  /* <= 0: quasistatic */
    double noise;               /* uniform perturbation of free nodes, in node spacings */
    uint64_t seed;
    double bc_stiffness;
    int32_t fixed_axis;
} symx_hex_block_params;

/* Newton settings exposed to callers (subset of NewtonSettings, solver_utils.h:173-258). */
typedef struct symx_newton_params {
    int32_t precision;               /* SYMX_HIP_FP64 or SYMX_HIP_MIXED */
    int32_t projection_mode;         /* 0 Newton, 1 ProjectedNewton, 2 ProjectOnDemand, 3 Progressive */
    int32_t use_mirroring;
    int32_t max_iterations;
    double residual_tolerance_abs;
    double residual_tolerance_rel;
    double cg_rel_tolerance;
    double cg_abs_tolerance;
    int32_t cg_max_iterations;
    double projection_eps;
    double step_tolerance;
    int32_t verbosity;               /* console output of the solve: 0 none, 1 summary, 2 one line per iteration, 3 full */
    int32_t linear_solver;           /* 0 BDPCG, 1 DirectLLT (solved to round-off on the device) */
} symx_newton_params;

typedef struct symx_solve_stats {
    int32_t result;              /* symx::SolverReturn */
    int32_t newton_iterations;
    int32_t cg_iterations;
    int32_t ls_bt_iterations;
    int32_t n_refinements;
    double final_energy;
    double final_residual;
    double first_residual;
    double seconds;
    double seconds_eval;         /* element evaluation (PGH) */
    double seconds_project;
    double seconds_assembly;
    double seconds_linear;
    double seconds_energy;       /* line-search energies */
    int32_t ls_trapezoid_accepts; /* line searches ended by the trapezoid estimate of the energy change */
    int32_t pad_;
} symx_solve_stats;

void symx_newton_params_default(symx_newton_params* p);
void symx_tet4_block_params_default(symx_tet4_block_params* p);

int symx_scene_tet4_block_create(const symx_tet4_block_params* p, const symx_newton_params* np, symx_scene* out);
void symx_scene_destroy(symx_scene s);
const char* symx_scene_last_error(void);

int symx_scene_n_vertices(symx_scene s, int32_t* n);
int symx_scene_n_elements(symx_scene s, int32_t* n);
/* connectivity [n_elem][nodes per element] */
int symx_scene_connectivity(symx_scene s, int32_t* conn);
/* positions (DoFs), previous positions x0, velocities v0, rest positions, masses */
int symx_scene_get_array(symx_scene s, const char* name, double* out);
int symx_scene_set_array(symx_scene s, const char* name, const double* in);
/* per-element log(mu) parameters (scenes created with log_mu_field = 1); also
 * refreshes the device copy of the scene's data, so direct symx_hip_* evaluations
 * see the new values */
int symx_scene_set_log_mu(symx_scene s, const double* log_mu);
int symx_scene_set_newton_params(symx_scene s, const symx_newton_params* np);
/* Newton solve of the current time step (does not advance time). */
int symx_scene_solve(symx_scene s, symx_solve_stats* st);
void symx_hex_block_params_default(symx_hex_block_params* p);
int symx_scene_hex_block_create(const symx_hex_block_params* p, const symx_newton_params* np, symx_scene* out);
int symx_scene_nodes_per_element(symx_scene s, int32_t* n);
/* per-element Young's moduli of a hexahedral scene (n_elements values) */
int symx_scene_young_moduli(symx_scene s, double* out);
void symx_sphere_params_default(symx_sphere_params* p);
int symx_scene_sphere_create(const symx_sphere_params* p, const symx_newton_params* np, symx_scene* out);
/* surface (contact) vertices of a sphere scene: *n receives the count; out may be NULL */
int symx_scene_contact_vertices(symx_scene s, int32_t* n, int32_t* out);

/* Adjoint of the quasistatic equilibrium in x (scenes with log_mu_field = 1):
 * solves H lambda = dL_dx on the unprojected Hessian at the current positions and
 * writes dL_dlog_mu[e] = -lambda^T d^2E/dx dlog_mu_e (n_elements values). Solved in
 * FP64 whatever the scene's precision; rel_tol is the PCG tolerance relative to |dL_dx|. */
typedef struct symx_adjoint_stats {
    int32_t cg_iterations;
    int32_t n_refinements;
    int32_t converged;
    int32_t found_indefiniteness;
    double cg_error;
    double energy;
    double seconds;
} symx_adjoint_stats;
int symx_scene_adjoint(symx_scene s, const double* dL_dx, double* dL_dlog_mu, double rel_tol, int32_t max_iter, symx_adjoint_stats* st);
/* Solve then advance: v = (x - x0)/dt, x0 = x. */
int symx_scene_step(symx_scene s, symx_solve_stats* st);
/* Device problem of the scene, for direct symx_hip_* calls. */
symx_hip_problem symx_scene_problem(symx_scene s);
/* Generated HIP source of potential i (NUL-terminated copy into buf). */
int symx_scene_source(symx_scene s, int32_t i, char* buf, int64_t cap, int64_t* len);
/* SSA operation counts of potential i: P, PG, PGH kernels. */
int symx_scene_op_counts(symx_scene s, int32_t i, int64_t* ops3);
double symx_scene_compile_seconds(symx_scene s);

/* ---- host-only helpers (no GPU): symbolic path checks ----------------------- */
/* Evaluate E, gradient and Hessian of the Neo-Hookean Tet4 element exactly as a
 * scene builds it (fem_integrator + neohookean energy density), with the host
 * reference evaluator of the expression graph. material: 0 NH, 1 stable NH
 * (then lambda, mu are read as E and nu). */
int symx_host_tet4_element(int32_t material, const double* x12, const double* X12, double lambda, double mu,
                           double* E, double* g12, double* H144);
/* Point-triangle signed distance (sign * (p - a).n) and its gradient and Hessian
 * w.r.t. (a, b, c, p): in[13] = a, b, c, p, sign; out[157] = value, grad[12], H[144]. */
int symx_host_point_triangle(const double* in13, double* out157);
/* Generate the HIP module of the Tet4 block scene's elastic potential and
 * compile it with hiprtc for gfx950 (no GPU needed); returns code size. */
int symx_host_jit_tet4(const char* cache_dir, int32_t material, int64_t* code_size, int64_t* n_ops_pgh);
/* Generate every potential module of a Tet4 block scene (adjoint kernels included
 * when log_mu_field = 1) and compile them with hiprtc for gfx950. */
int symx_host_scene_jit(const symx_tet4_block_params* p, const char* cache_dir, int32_t* n_potentials, int64_t* code_bytes,
                        int64_t* ops_adjoint);
/* Sphere mesh of symx_scene_sphere_create on the host: counts, and (when not NULL)
 * rest positions [nv][3], tets [ne][4], contact vertices [nc]. */
int symx_host_sphere_mesh(const symx_sphere_params* p, int32_t* nv, int32_t* ne, int32_t* nc, double* X, int32_t* tets, int32_t* contact);
/* Generate and hiprtc-compile every potential module of a hexahedral block scene;
 * *n_structured counts the potentials compiled as structured FEM kernels. */
int symx_host_hex_jit(const symx_hex_block_params* p, const char* cache_dir, int32_t* n_potentials, int64_t* code_bytes, int32_t* n_structured);
/* E, gradient [n] and Hessian [n][n] (DoF order) of element e of a hexahedral block's
 * elastic potential evaluated on the host from the generated expressions: structured
 * (deferred Jacobian, the structured kernel's contraction) or generic (expanded
 * expression). *was_structured reports the path taken. */
int symx_host_hex_element(const symx_hex_block_params* p, int32_t structured, int32_t e, double* E, double* g, double* H, int32_t* was_structured);
/* Evaluate the generated adjoint outputs d(lambda . dE_e/dx)/dlog_mu_e of element e of
 * a log_mu_field scene on the host (symbolic graph, kernel's gather rules); lambda
 * has one value per DoF. */
/* Generated E, gradient (n) and Hessian (n x n, row-major) of element e of potential
 * `potential` of a Tet4 block scene, evaluated on the host from the expressions the element
 * kernels are generated from: with chain_rule != 0 the derivatives go through the linear cut
 * of the expression where that is chosen (*was_factorised = 1), else they are the flat
 * symbolic derivatives. *n_dofs receives n (g and H must hold 36 and 1296 values). */
int symx_host_scene_element(const symx_tet4_block_params* p, int32_t chain_rule, int32_t potential, int32_t e, double* E, double* g,
                            double* H, int32_t* n_dofs, int32_t* was_factorised);
int symx_host_scene_adjoint_element(const symx_tet4_block_params* p, const double* log_mu, int32_t e, const double* lambda, double* out);

#ifdef __cplusplus
}
#endif
#endif
