/*
 * symx_hip.h - C ABI of the MI355X (gfx950) evaluation and solve backend.
 *
 * Everything device-side is owned by this library: JIT code objects, device
 * DoF vectors, element outputs, the assembled block matrix and the PCG state.
 * The C++ front end (GlobalPotential / NewtonsMethod / Compiled<double>) and
 * the ctypes test harness call only these entry points. All functions return
 * SYMX_HIP_OK (0) or an error code; the message is available from
 * symx_hip_last_error(). Nothing here throws, and no torch type appears in a
 * signature. Calls on one context are not thread safe; use one context per
 * host thread. Calls are synchronous at return.
 *
 * Each entry point names the SymX code it replaces (paths relative to the
 * SymX repository root).
 */
#ifndef SYMX_HIP_H
#define SYMX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
enum {
    SYMX_HIP_OK = 0,
    SYMX_HIP_ERR_ARG = 1,       /* invalid argument / state */
    SYMX_HIP_ERR_HIP = 2,       /* HIP runtime error */
    SYMX_HIP_ERR_JIT = 3,       /* hiprtc compilation or module load failed */
    SYMX_HIP_ERR_NO_DEVICE = 4, /* no GPU visible */
    SYMX_HIP_ERR_ALLOC = 5      /* device allocation failed */
};

/* ---- enums mirrored from SymX ---------------------------------------------- */
/* Precision of element Hessian storage, block matrix and PCG. */
enum { SYMX_HIP_FP64 = 0, SYMX_HIP_MIXED = 1 };
/* What a potential evaluation produces (SecondOrderCompiledGlobal.h:27-31). */
enum { SYMX_HIP_EVAL_P = 0, SYMX_HIP_EVAL_PG = 1, SYMX_HIP_EVAL_PGH = 2 };
/* PGH with the element Hessians stored as floats (each potential's kernel_PGH_f32): the
 * storage of SYMX_HESS_STORAGE_FLOAT=float moved from the block matrix
 * (ElementHessians.cpp:236-249) to the element Hessians; projection and assembly then
 * read floats, work and sum in FP64, and store in their own precision. */
enum { SYMX_HIP_EVAL_PGH_F32 = 3 };
/* Named device vectors of a problem, for the Newton vector operations. */
enum {
    SYMX_HIP_VEC_DOFS = 0,    /* current DoFs (what element kernels read) */
    SYMX_HIP_VEC_GRAD = 1,    /* assembled gradient */
    SYMX_HIP_VEC_DU = 2,      /* Newton step */
    SYMX_HIP_VEC_DOFS_LS = 3, /* DoFs saved at line-search start */
    SYMX_HIP_VEC_RHS = 4      /* linear-system right-hand side */
};

#define SYMX_HIP_MAX_MAPS 128
#define SYMX_HIP_BLOCK 256 /* threads per workgroup of every generated element kernel */

/* Argument block of every generated element kernel (one struct by value).
 * Gather contract (CompiledInLoop_run.h:170-190, fetch_blocks.h:11-31):
 *   symbol f+k of map m  =  maps[m][ conn[c_m * n_elem + e] * s_m + k ]   (c_m >= 0)
 *                        =  maps[m][ k ]                                  (fixed map)
 * Element outputs: E_partial[blockIdx.x] is the FP64 sum over the block's
 * elements (fixed order); G is [n_elem][n_dof]; H holds, per element, the
 * upper 3x3 blocks (a <= b, row-major order of (a,b)) with 9 row-major values
 * each, as doubles, or as floats for the kernels emitted for FP32 Hessian
 * storage (name suffix "_f32"), in records of h_stride values (the potential's
 * symx_hip_potential_desc; pad values are written as zeros). Inactive elements
 * (active[e] == 0) write zeros and have energy 0.
 * e_mode selects what happens to the element energies E_e:
 *   SYMX_HIP_EMODE_SUM    E_partial gets the sum of E_e;
 *   SYMX_HIP_EMODE_STORE  the same, and E_elem[e] = E_e for every e < n_elem;
 *   SYMX_HIP_EMODE_DELTA  E_partial gets the sum of (E_e - E_elem[e]). */
enum { SYMX_HIP_EMODE_SUM = 0, SYMX_HIP_EMODE_STORE = 1, SYMX_HIP_EMODE_DELTA = 2 };
typedef struct symx_hip_eval_params {
    const int32_t* conn;
    const uint8_t* active;
    const double* maps[SYMX_HIP_MAX_MAPS];
    double* E_partial;
    double* G;
    double* H;
    uint8_t* C_out;
    double* E_elem;
    int32_t n_elem;
    int32_t e_mode;
} symx_hip_eval_params;

/* ---- context ---------------------------------------------------------------- */
typedef struct symx_hip_ctx_s* symx_hip_ctx;
typedef struct symx_hip_module_s* symx_hip_module;
typedef struct symx_hip_problem_s* symx_hip_problem;

/* Number of visible GPUs (0 on a host without one). */
int symx_hip_device_count(int* n);
int symx_hip_ctx_create(int device, symx_hip_ctx* out);
void symx_hip_ctx_destroy(symx_hip_ctx ctx);
/* Last error message of ctx, or of the most recent context-less call if ctx is NULL. */
const char* symx_hip_last_error(symx_hip_ctx ctx);

/* ---- JIT (replaces symx/src/compile/Compilation.cpp:59-226) ---------------- */
/* Device-side prelude every generated element module starts with (the
 * argument struct above plus helpers); the emitter prepends it verbatim. */
const char* symx_hip_eval_prelude(void);
/* Compile HIP source with hiprtc for `arch` (NULL = "gfx950") into a code
 * object cached under cache_dir as <sha256>.hsaco (key = SHA-256 of source,
 * hiprtc version, arch and options; the file holds the code's size and SHA-256
 * ahead of it, and a file that does not match them is compiled again). Needs
 * no GPU. key_out (>= 65 bytes) receives the key; *was_cached tells whether the
 * cache hit. Replaces the g++ + dlopen JIT (Compilation.cpp:84-127) and its
 * SHA-256 cache check (:128-226). */
int symx_hip_jit_compile(const char* source, const char* arch, const char* cache_dir,
                         char* key_out, int* was_cached, size_t* code_size);
/* Compile (or fetch from cache) and load onto the context's device. */
int symx_hip_module_load(symx_hip_ctx ctx, const char* source, const char* cache_dir, symx_hip_module* out);
void symx_hip_module_unload(symx_hip_module mod);
/* Launch generated kernel `name` of `mod` over n_elem elements with the given
 * argument block (E_partial must hold ceil(n_elem/256) doubles). */
int symx_hip_module_launch_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name,
                                const symx_hip_eval_params* params);

/* ---- problem: device-resident DoFs + potentials ----------------------------
 * Replaces SecondOrderCompiledGlobal (SecondOrderCompiledGlobal.cpp:9-142),
 * SecondOrderCompiledPotential's evaluation callbacks
 * (SecondOrderCompiledPotential.cpp:94-197), Assembly/ElementHessians
 * (ElementHessians.cpp:5-293) and the BlockedSparseMatrix used by
 * NewtonsMethod (NewtonsMethod.cpp:143-146, 388-457). */
int symx_hip_problem_create(symx_hip_ctx ctx, int32_t n_dofs, symx_hip_problem* out);
void symx_hip_problem_destroy(symx_hip_problem prob);

/* One data map of a potential (data_maps.h:86-105). */
typedef struct symx_hip_map_desc {
    int32_t is_dof;        /* 1: the data is the DoF vector at dof_offset */
    int32_t dof_offset;    /* offset of this DoF set in the flat DoF vector */
    const double* host;    /* non-DoF data (copied at creation / refresh) */
    int64_t n_values;      /* number of doubles in host data */
} symx_hip_map_desc;

typedef struct symx_hip_potential_desc {
    symx_hip_module module;
    const char* kernel_P;     /* energy                              */
    const char* kernel_PG;    /* energy + gradient                   */
    const char* kernel_PGH;   /* energy + gradient + upper Hessian    */
    const char* kernel_C;     /* activity condition (NULL if none)    */
    const char* kernel_PGH_f32; /* PGH storing the Hessian as floats (NULL if none: no SYMX_HIP_EVAL_PGH_F32) */
    int32_t n_elem;
    int32_t conn_stride;
    const int32_t* conn;      /* host, row-major [n_elem][conn_stride] */
    int32_t n_dof_blocks;     /* DoF blocks (3 DoFs each) per element */
    const int32_t* dof_conn_col;   /* [n_dof_blocks] connectivity column of block i */
    const int32_t* dof_set_offset; /* [n_dof_blocks] DoF offset of block i's set */
    int32_t n_maps;
    const symx_hip_map_desc* maps;
    /* Elements per 256-thread workgroup of the P/PG/PGH kernels and per energy partial:
     * 0 or 256 for one element per thread (generated element kernels), 4 for one element
     * per wavefront (structured FEM kernels). The C kernel always runs one per thread. */
    int32_t elems_per_group;
    /* Values per element Hessian record on the device: 0 for n_ub * 9 (upper 3x3 blocks,
     * contiguous). Larger (a multiple of 3) when the PGH kernels pad each record, e.g. 96
     * for a 4-node element's 90 values, so that every 128-byte chunk the kernel stores is a
     * whole aligned line; the pad stays out of projection and assembly. */
    int32_t h_stride;
} symx_hip_potential_desc;

int symx_hip_problem_add_potential(symx_hip_problem prob, const symx_hip_potential_desc* desc, int32_t* pot_id);
/* Refresh the device copy of a non-DoF map (e.g. x0 or v0 between time steps). */
int symx_hip_problem_set_map_data(symx_hip_problem prob, int32_t pot_id, int32_t map_id, const double* host, int64_t n_values);
/* Build the assembly topology (gradient scatter lists, block-matrix pattern,
 * slot-sorted contribution lists). Must be called after the last
 * add_potential and before any evaluation. */
int symx_hip_problem_finalize(symx_hip_problem prob);

/* DoF transfer (GlobalPotential.cpp:111-144). */
int symx_hip_problem_set_dofs(symx_hip_problem prob, const double* host);
int symx_hip_problem_get_dofs(symx_hip_problem prob, double* host);
/* Copy one named device vector to / from host memory. */
int symx_hip_problem_download(symx_hip_problem prob, int32_t vec, double* host);
int symx_hip_problem_upload(symx_hip_problem prob, int32_t vec, const double* host);
/* Device pointer of a named vector (valid until the problem is destroyed). */
int symx_hip_problem_device_ptr(symx_hip_problem prob, int32_t vec, double** dptr);

/* Evaluate all potentials at the current DoFs (evaluate_P / evaluate_P__dP_du /
 * evaluate_P__dP_du__local_d2P_du2, SecondOrderCompiledGlobal.cpp:72-142).
 * The energy is returned in *E (FP64, deterministic order); the gradient is
 * assembled into VEC_GRAD for PG and PGH; PGH keeps element Hessians on the
 * device for project / assemble. */
int symx_hip_problem_eval(symx_hip_problem prob, int32_t mode, double* E);
/* Energy change since the last EVAL_PG / EVAL_PGH: *dE = sum over potentials and
 * elements, in a fixed order, of (E_e(current DoFs) - E_e(DoFs of that evaluation)).
 * This is the quantity the Armijo test of the line search compares with the expected
 * decrease (E1 < E0 + beta s du.g, NewtonsMethod.cpp:565-591). Forming it element by
 * element keeps its round-off at the size of the element energies instead of the
 * total energy's, whose FP64 resolution on millions of elements is of the order of the
 * decrease that a Newton step near a 1e-6 relative residual predicts. */
int symx_hip_problem_energy_change(symx_hip_problem prob, double* dE);
/* *out = g(x) . vec, with g(x) the gradient of the total energy at the current DoFs
 * (evaluate_P__dP_du, SecondOrderCompiledGlobal.cpp:90-116) assembled into a scratch
 * vector: VEC_GRAD, the element Hessians and the reference energies of
 * symx_hip_problem_energy_change are left as they are. The line search uses it for the
 * trapezoid estimate E(x0 + d) - E(x0) ~ (g(x0) + g(x0 + d)) . d / 2 of steps too short
 * for energy differences to resolve. */
int symx_hip_problem_gradient_dot(symx_hip_problem prob, int32_t vec, double* out);
/* Number of element Hessians produced by the last PGH evaluation. */
int symx_hip_problem_n_element_hessians(symx_hip_problem prob, int64_t* n);
/* Copy element outputs of potential pot_id to host (tests): G [n_elem][n_dof],
 * H [n_elem][n_ub*9] (without the records' pad); either pointer may be NULL. */
int symx_hip_problem_element_outputs(symx_hip_problem prob, int32_t pot_id, double* G, double* H);

/* Project every element Hessian to PD (project_to_PD.cpp:13-82 via
 * ElementHessians::project_to_PD_inplace__all, ElementHessians.cpp:48-67):
 * eigenvalues < eps become eps (clamp) or -lambda (mirror). *n_changed
 * receives the number of matrices that were rebuilt; with n_changed NULL the
 * call only enqueues the work on the context's stream and returns at once. */
int symx_hip_problem_project(symx_hip_problem prob, double eps, int32_t mirror, int64_t* n_changed);
/* Progressive projection step (NewtonsMethod.cpp:306-330 with
 * ElementHessians::project_to_PD_for_update__selectively, ElementHessians.cpp:68-165):
 * a DoF block is active when max |VEC_GRAD| over its 3 entries >= threshold (0: all
 * blocks); every element touching an active block that has not been projected since
 * the last EVAL_PGH is projected and remembered. Outputs (each may be NULL): matrices
 * rebuilt, elements newly projected, active blocks. Assemble afterwards. */
int symx_hip_problem_project_selective(symx_hip_problem prob, double eps, int32_t mirror, double threshold, int64_t* n_changed,
                                       int64_t* n_newly_projected, int64_t* n_active_blocks);
/* Assemble element Hessians into the device block matrix in FP64 or FP32
 * storage (ElementHessians::assemble_global, ElementHessians.cpp:224-256) and
 * prepare the block-Jacobi preconditioner (BlockedSparseMatrix.h:1147-1214).
 * Sums run in a fixed order: results are bitwise repeatable. The work is
 * enqueued on the context's stream (later calls on the problem are ordered
 * after it). */
int symx_hip_problem_assemble(symx_hip_problem prob, int32_t precision);
/* Export the assembled matrix as block-CSR (row offsets [n_blocks+1], block
 * columns [nnzb], values [nnzb][9] row-major, FP64) - used by tests. Any
 * pointer may be NULL to query *nnzb only. */
int symx_hip_problem_export_bsr(symx_hip_problem prob, int64_t* nnzb, int64_t* row_ptr, int32_t* cols, double* vals);

/* PCG result (solve_pcg.h:21-30). */
typedef struct symx_hip_pcg_info {
    int32_t n_iterations;
    int32_t converged;
    int32_t found_indefiniteness;
    int32_t n_refinements; /* outer FP64 refinement steps (mixed precision) */
    double error;          /* ||r|| / ||b|| of the returned solution */
} symx_hip_pcg_info;

/* Solve A du = rhs with block-Jacobi PCG from du = 0, with the stopping rules
 * of bsm::solve_pcg (solve_pcg.h:83-232): exit if ||b||^2 < abs_tol^2; stop
 * when ||r||/||b|| < abs_tol or (||r||/||b||)/err0 < rel_tol; stop on
 * p.Ap <= 0 when stop_on_indef. Precision FP64 runs exactly that iteration in
 * FP64; MIXED runs it in FP32 on the FP32 matrix inside FP64 iterative
 * refinement whose outer residual b - A x is FP64, stopping on the same rules. */
int symx_hip_problem_solve(symx_hip_problem prob, int32_t precision, double abs_tol, double rel_tol,
                           int32_t max_iter, int32_t stop_on_indef, symx_hip_pcg_info* info);

/* Newton vector operations on named device vectors (NewtonsMethod.cpp:105-218, 459-641). */
int symx_hip_problem_max_abs(symx_hip_problem prob, int32_t vec, double* out);
int symx_hip_problem_dot(symx_hip_problem prob, int32_t a, int32_t b, double* out);
/* dst = a + s * b */
int symx_hip_problem_axpy(symx_hip_problem prob, int32_t dst, int32_t a, double s, int32_t b);
/* dst = s * a */
int symx_hip_problem_scale(symx_hip_problem prob, int32_t dst, int32_t a, double s);

/* Evaluate generated batch kernel `name` (emitted for Compiled<double>) on n
 * input vectors: host_in [n][n_in] -> host_out [n][n_out]. Replaces the
 * single-call Compiled<T>::run (Compiled.cpp:117-179) with a batched launch. */
int symx_hip_batch_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const double* host_in, int64_t n,
                        int32_t n_in, double* host_out, int32_t n_out);

/* ---- adjoint of a quasistatic equilibrium (no counterpart in SymX; composed
 * from its pieces: unprojected Hessian, PCG, mixed partial kernels) ----------
 * Solve H lambda = rhs, where H is assembled here, without PD projection, from the
 * element Hessians of the last EVAL_PGH (the equilibrium). rhs (host, n_dofs) is
 * dL/dx; lambda stays on the device in VEC_DU and is copied to host_lambda when
 * it is not NULL. Stopping rule of symx_hip_problem_solve with abs_tol = 0. */
int symx_hip_problem_adjoint_solve(symx_hip_problem prob, int32_t precision, double rel_tol, int32_t max_iter,
                                   const double* host_rhs, double* host_lambda, symx_hip_pcg_info* info);
/* Second argument of generated adjoint kernels, which have the signature
 * (const symx_hip_eval_params p, const symx_hip_adjoint_params a). lambda[m] points
 * into the adjoint vector (VEC_DU) at the DoF offset of data map m (NULL for
 * non-DoF maps); the kernel writes out[e * n_out + j] = d(lambda . dE_e/dx)/dp_j. */
typedef struct symx_hip_adjoint_params {
    const double* lambda[SYMX_HIP_MAX_MAPS];
    double* out;
    int32_t n_out;
    int32_t pad_;
} symx_hip_adjoint_params;

/* Launch the adjoint kernel `name` (in the potential's own module) over the
 * elements of potential pot_id with lambda = VEC_DU; out [n_elem][n_out] is copied
 * to host_out. */
int symx_hip_problem_launch_adjoint(symx_hip_problem prob, int32_t pot_id, const char* name, int32_t n_out, double* host_out);

/* Step filter against a plane (the max_allowed_step callback of SymX's line search,
 * solver_utils.h:74 and NewtonsMethod.cpp:494-506, for contact with x[axis] = plane):
 * set_plane_filter stores the DoF blocks (vertices) to watch; plane_max_step returns
 * t = min(1, min over watched blocks with du[axis] < 0 of eta * (x[axis] - plane) / -du[axis])
 * from VEC_DOFS and VEC_DU, so a step of t * du keeps every watched vertex on its side. */
int symx_hip_problem_set_plane_filter(symx_hip_problem prob, const int32_t* blocks, int32_t n_blocks, int32_t axis, double plane);
int symx_hip_problem_plane_max_step(symx_hip_problem prob, double eta, double* t_max);

/* Stateless per-element evaluation (CompiledInLoop, CompiledInLoop_run.h:66-472):
 * uploads the connectivity conn [n_elem][conn_stride] (host, row-major) and the data
 * maps (host pointers, maps[m].n_values values each; is_dof is ignored), launches the
 * kernel `name` of `mod` (signature (const symx_hip_eval_params p, const
 * symx_hip_adjoint_params a), writing a.out[e * n_out + j]) and copies the outputs to
 * host_out [n_elem][n_out]. active (host, may be NULL) holds one flag per element;
 * inactive elements get zeros. Device buffers are freed before returning. */
int symx_hip_loop_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const int32_t* conn, int32_t n_elem, int32_t conn_stride,
                       const symx_hip_map_desc* maps, int32_t n_maps, const uint8_t* active, int32_t n_out, double* host_out);

/* Allocation-free variant of symx_hip_loop_eval. No counterpart in SymX: it is for callers
 * that cannot allocate device memory (code that must not call hipMalloc, such as tight
 * serving loops or harnesses that forbid allocation). Inputs and outputs are DEVICE
 * pointers owned by the caller: d_conn [n_elem][conn_stride] row-major, d_maps[m] (host
 * array of n_maps device pointers), d_active (may be NULL; one flag per element) and d_out
 * [n_elem][n_out]. The connectivity is transposed into a per-GPU scratch array of
 * SYMX_HIP_LOOP_SCRATCH_ENTRIES int32 values (16 MiB) that the library holds from the first
 * symx_hip_ctx_create on the GPU until the process ends; calls with n_elem * conn_stride
 * above that limit fail with SYMX_HIP_ERR_ARG. The call makes no hipMalloc / hipFree. Every
 * call on one GPU shares the scratch array: calls that overlap on one GPU (several contexts,
 * streams or host threads) race. Synchronous at return, like the other entry points. */
#define SYMX_HIP_LOOP_SCRATCH_ENTRIES (1 << 22)
int symx_hip_loop_eval_noalloc(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const int32_t* d_conn, int32_t n_elem,
                               int32_t conn_stride, const double* const* d_maps, int32_t n_maps, const uint8_t* d_active, int32_t n_out,
                               double* d_out);

/* Kernel timing with HIP events on the context's stream (the stream every
 * kernel of this library runs on): average duration in ms of `reps` launches
 * of one phase. phase: 0 element kernel E+grad+Hessian of potential `pot`,
 * 1 PD projection of `pot`'s element Hessians (in place, on a scratch copy),
 * 2 block-matrix assembly (arg = precision), 3 one PCG iteration (spmv + update,
 * arg = precision), 4 energy-only element kernel of `pot`. Needs a prior
 * EVAL_PGH (phases 1-3) and assemble (phase 3). No counterpart in SymX: its
 * Logger::ScopedTimer (Logger.h:17-29) times phases on the host. */
int symx_hip_problem_profile(symx_hip_problem prob, int32_t phase, int32_t pot, int32_t arg, int32_t reps, double* avg_ms);

/* ---- standalone kernels on caller-provided device memory (tests) ---------- */
/* In-place PD projection of n_mats dense symmetric row-major n x n FP64 matrices,
 * n in {3, 6, 9, 12, 24, 30, 81} (project_to_PD.cpp:13-82). changed (nullable, device)
 * gets 1 per rebuilt matrix. */
int symx_hip_project_dense(symx_hip_ctx ctx, double* d_mats, int64_t n_mats, int32_t n, double eps,
                           int32_t mirror, uint8_t* d_changed);

#ifdef __cplusplus
}
#endif
#endif
