// Internal state of the C ABI (include/symx_hip.h). Host-side only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/symx_hip.h"

namespace symx_hip
{
    // Device scalars of one PCG run; lives in device memory, mirrored to pinned host memory.
    struct PcgState
    {
        double bb;      // ||b||^2
        double rz;      // r.z of the current iterate
        double pAp;
        double alpha;
        double beta;
        double rr;      // ||r||^2
        double err;     // ||r|| / ||b||
        double err0;
        double abs_tol;
        double rel_tol;
        int32_t it;
        int32_t max_iter;
        int32_t done;
        int32_t converged;
        int32_t indef;
        int32_t stop_on_indef;
    };

    // Sliced block-CSR (SELL-64 of 3x3 blocks): block rows are placed at positions
    // grouped in slices of 64; slice s stores width_s block columns for each of its
    // positions, the diagonal block first. Slot (s, j, lane) = slice_ptr[s] + 64 j +
    // lane; the 64 blocks of one (s, j) form a group of 576 values starting at
    // (slice_ptr[s] / 64 + j) * 576 (sell_val_index): values 0..7 of a block (row-major)
    // in runs of W = SellRun<T>::width consecutive values per lane, the runs of the 64
    // lanes side by side, then value 8 of every lane. A load of one run is then W values
    // per lane and 64 W consecutive values per wave (16-byte loads for FP32). (A layout
    // with each block row as one 3-value load per lane made the FP64 SpMV 25% slower and
    // the FP32 one no faster.)
    template <typename T>
    struct SellRun
    {
        static constexpr int width = sizeof(T) == 4 ? 4 : 1;
    };
    template <typename T>
    __host__ __device__ constexpr int64_t sell_val_index(int64_t slot, int q)
    {
        constexpr int W = SellRun<T>::width;
        const int64_t g = (slot >> 6) * 576;
        const int lane = (int)(slot & 63);
        return q < 8 ? g + (q / W) * 64 * W + lane * W + q % W : g + 512 + lane;
    }
    // Position = row unless sorting the rows by length inside windows of kSellSigma
    // rows removes enough padding (SELL-C-sigma); then row_at[pos] names the row.
    constexpr int32_t kSellSigma = 1024;
    // SpMV and PCG-update workgroup size. The grid (pcg_grid) is a fixed set of reduction
    // workgroups, each owning a contiguous chunk of positions; wave v of a workgroup takes
    // slices v, v + 8, ... of its chunk. 512 threads keep 8 waves per SIMD in flight; a
    // streaming read reaches 7.3 TB/s with 8 waves per SIMD and 5.5 TB/s with 4.
    constexpr int kSpmvThreads = 512;
    // Workgroups of the PCG kernels for n_rows block rows.
    int pcg_grid(int32_t n_rows);
    struct SellPattern
    {
        int32_t n_rows = 0;
        int32_t n_slices = 0;
        int64_t n_slots = 0;
        int64_t nnzb = 0;                 // real (non padding) blocks
        std::vector<int64_t> h_slice_ptr; // [n_slices + 1], multiples of 64
        std::vector<int32_t> h_col;       // [n_slots]
        std::vector<int32_t> h_row_len;   // [n_rows]
        std::vector<int32_t> h_row_at;    // [n_rows] row at each position; empty = identity
        std::vector<int32_t> h_pos;       // [n_rows] position of each row; empty = identity
        int64_t* d_slice_ptr = nullptr;
        int32_t* d_col = nullptr;         // column ids, or
        int16_t* d_dcol = nullptr;        // column - row per slot when every offset fits 16 bits
        int32_t* d_row_at = nullptr;      // null = identity
        int32_t pos_of(int32_t r) const { return h_pos.empty() ? r : h_pos[r]; }
    };

    void set_thread_error(const std::string& msg);
}

struct symx_hip_ctx_s
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // Reduction scratch: per-workgroup partials and the last-block ticket.
    double* d_partials = nullptr;
    size_t partials_cap = 0;
    unsigned int* d_counters = nullptr;  // small device counters (length of the projection fallback list)
    double* d_scalars = nullptr;   // small device scalar area
    double* h_pinned = nullptr;    // pinned host scalars (64 doubles)
    symx_hip::PcgState* d_pcg = nullptr;
    symx_hip::PcgState* h_pcg = nullptr;  // two pinned copies: the PCG batch pipeline reads one while the other is in flight
    hipEvent_t pcg_ev[2] = { nullptr, nullptr };
    int32_t* d_loop_scratch = nullptr;  // device address of this device's loop scratch array
    char* h_stage[2] = { nullptr, nullptr };  // pinned staging of large host transfers (allocated on first use)
    hipEvent_t stage_ev[2] = { nullptr, nullptr };
};

struct symx_hip_module_s
{
    symx_hip_ctx ctx = nullptr;
    hipModule_t mod = nullptr;
    std::unordered_map<std::string, hipFunction_t> fns;
};

namespace symx_hip
{
    struct DevMap
    {
        int32_t is_dof = 0;
        int32_t dof_offset = 0;
        double* d_data = nullptr;  // device copy of non-DoF data
        int64_t n_values = 0;
        bool owner = true;           // false: d_data belongs to an earlier map of the same host array
        const double* host = nullptr;
    };

    struct DevPotential
    {
        symx_hip_module module = nullptr;
        std::string kP, kPG, kPGH, kC;
        std::string kPGH32;                 // PGH with FP32 Hessian storage (empty: none)
        int32_t n_elem = 0;
        int32_t stride = 0;
        std::vector<int32_t> h_conn;        // row-major copy (topology)
        int32_t* d_conn = nullptr;          // [stride][n_elem]
        int32_t nb = 0;                     // DoF blocks per element
        int32_t n_ub = 0;                   // upper 3x3 blocks per element
        std::vector<int32_t> dof_col, dof_off;
        std::vector<DevMap> maps;
        uint8_t* d_active = nullptr;
        int64_t g_base = 0;                 // first gradient block (3 doubles)
        int32_t h_stride = 0;               // values per element Hessian record (>= n_ub * 9)
        int64_t h_off = 0;                  // first value of the potential's Hessian records
        int64_t part_base = 0;              // first energy partial
        int64_t e_base = 0;                 // first element in problem-wide element numbering
        int32_t n_groups = 0;               // workgroups of the element kernel
        int32_t elems_per_group = SYMX_HIP_BLOCK;
    };
}

struct symx_hip_problem_s
{
    symx_hip_ctx ctx = nullptr;
    int32_t n_dofs = 0;
    int32_t nb = 0;
    double* d_vec[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
    std::vector<symx_hip::DevPotential> pots;
    bool finalized = false;

    // element outputs
    int64_t total_g = 0, total_h = 0, total_parts = 0;
    double* d_G = nullptr;
    double* d_H = nullptr;
    double* d_Epart = nullptr;
    double* d_Eelem = nullptr;    // [total_elem] element energies of the last EVAL_PG / EVAL_PGH
    bool e_ref_valid = false;
    uint8_t* d_flags = nullptr;   // per-element scratch of the projection
    uint32_t* d_plist = nullptr;  // per-element scratch: elements left to the fallback projection
    bool h_valid = false;
    bool h_f32 = false;           // element Hessians of the last PGH evaluation are stored as floats
    // max|grad| computed with the gradient by EVAL_PG / EVAL_PGH (read back with the energy),
    // valid until the gradient vector is written otherwise; unused once a raw device
    // pointer to the gradient has been handed out
    bool grad_max_valid = false;
    bool grad_ptr_exposed = false;
    double grad_max = 0.0;
    int64_t total_elem = 0;
    // progressive projection: elements projected since the last EVAL_PGH, and the
    // selection of the current call (problem-wide element numbering; allocated on first use)
    uint8_t* d_projected = nullptr;
    uint8_t* d_select = nullptr;
    // plane step filter
    int32_t* d_filter_blocks = nullptr;
    int32_t n_filter_blocks = 0;
    int32_t filter_axis = 2;
    double filter_plane = 0.0;

    // gradient scatter: CSR of gradient-block ids by DoF block, the blocks in the assembly
    // traversal order (d_grow[t] = DoF block of list t)
    uint32_t* d_goff = nullptr;
    uint32_t* d_glist = nullptr;
    int32_t* d_grow = nullptr;

    // block matrix
    symx_hip::SellPattern sell;
    uint32_t* d_aslot = nullptr;  // [n_assemble] assembled slots in traversal order
    uint32_t* d_mirror = nullptr; // [n_assemble] slot that receives the transposed block, ~0u for none; null: no mirroring
    int64_t n_assemble = 0;
    uint32_t* d_coff = nullptr;   // [n_slots + 1] contribution lists in that order
    uint32_t* d_clist = nullptr;  // (offset of the Hessian block's first value / 3) * 2 + transpose flag
    double* d_val64 = nullptr;
    float* d_val32 = nullptr;
    double* d_dinv64 = nullptr;
    float* d_dinv32 = nullptr;
    int32_t assembled_precision = -1;

    // PCG work vectors
    double *d_x = nullptr, *d_r = nullptr, *d_z = nullptr, *d_p0 = nullptr, *d_p1 = nullptr, *d_q = nullptr;
    float *f_x = nullptr, *f_r = nullptr, *f_z = nullptr, *f_p0 = nullptr, *f_p1 = nullptr, *f_q = nullptr, *f_b = nullptr;
};

// Launch wrappers implemented in kernels.hip ------------------------------------
namespace symx_hip
{
    // Fixed-order sum of n doubles into *d_out (single workgroup).
    void launch_sum(hipStream_t s, const double* d_in, int64_t n, double* d_out);
    void launch_grad_assemble(hipStream_t s, const uint32_t* off, const uint32_t* list, const int32_t* rows, const double* G, double* grad,
                              int32_t nb);
    // flags: n_mats bytes of scratch (4-block elements that are not translation invariant)
    // Elements of the active DoF blocks (max |grad| of the block >= threshold) that are not
    // projected yet get select[e] = 1; the number of active blocks is added to n_active.
    struct PotRanges
    {
        int32_t n = 0;
        int32_t nb[16];
        int64_t g_base[16];
        int64_t g_end[16];
        int64_t e_base[16];
        const uint8_t* active[16];  // activity flags of conditional potentials (NULL: all active)
    };
    void launch_ppn_select(hipStream_t s, const uint32_t* off, const uint32_t* list, const int32_t* rows, const double* grad, int32_t nblocks,
                           double threshold,
                           const PotRanges& pr, const uint8_t* projected, uint8_t* select, unsigned long long* n_active);
    // projected |= select, select = 0; adds the number of newly projected elements to n_new.
    void launch_ppn_commit(hipStream_t s, uint8_t* projected, uint8_t* select, int64_t n, unsigned long long* n_new);
    void launch_fill_u8(hipStream_t s, uint8_t* p, int64_t n);
    // out (one uint64, set to the bits of 1.0 first) = min over blocks of the plane step bound
    void launch_plane_max_step(hipStream_t s, const int32_t* blocks, int32_t n, const double* x, const double* du, int32_t axis, double plane,
                               double eta, unsigned long long* out);
    // select (may be NULL): only elements with select[e] != 0 are projected.
    // flags, list: per-element scratch (n_mats entries); n_list: one device counter.
    int launch_project_ub(hipStream_t s, double* H, int64_t n_mats, int32_t nb, int64_t rs, double eps, int32_t mirror, uint8_t* flags,
                          uint32_t* list, unsigned int* n_list, const uint8_t* select, unsigned long long* d_count);
    int launch_project_ub(hipStream_t s, float* H, int64_t n_mats, int32_t nb, int64_t rs, double eps, int32_t mirror, uint8_t* flags,
                          uint32_t* list, unsigned int* n_list, const uint8_t* select, unsigned long long* d_count);
    int launch_project_dense(hipStream_t s, double* mats, int64_t n_mats, int32_t n, double eps, int32_t mirror, uint8_t* changed);
    // n entries of aslot/coff; mirror (may be NULL): slot that also receives the transposed block
    // H32 (may be NULL): FP32 element-Hessian storage, read instead of H; sums run in FP64
    void launch_bsr_assemble(hipStream_t s, int64_t n, const uint32_t* aslot, const uint32_t* mirror, const uint32_t* coff,
                             const uint32_t* clist, const double* H, const float* H32, double* v64, float* v32);
    void launch_dinv(hipStream_t s, const SellPattern& P, const double* v64, const float* v32, double* dinv64, float* dinv32);
    // PCG on the sell matrix (iteration k is 1-based). partials: >= 5 * reduction_groups(n_rows) doubles.
    void launch_pcg_init64(hipStream_t s, const SellPattern& P, const double* dinv, const double* b, double* x, double* r, double* z,
                           double* p0, double* partials);
    void launch_pcg_iter64(hipStream_t s, const SellPattern& P, int32_t k, const double* val, const double* dinv, double* x, double* r,
                           double* z, const double* p_old, double* p_new, double* q, symx_hip::PcgState* st, double* partials);
    void launch_pcg_init32(hipStream_t s, const SellPattern& P, const float* dinv, const float* b, float* x, float* r, float* z,
                           float* p0, double* partials);
    void launch_pcg_iter32(hipStream_t s, const SellPattern& P, int32_t k, const float* val, const float* dinv, float* x, float* r,
                           float* z, const float* p_old, float* p_new, float* q, symx_hip::PcgState* st, double* partials);
    // Mixed-precision refinement: r = b - A x (FP64 vectors, FP32 matrix, FP64 arithmetic), r32 = (float) r, *rr_out = ||r||^2.
    void launch_residual_mixed(hipStream_t s, const SellPattern& P, const float* val, const double* b, const double* x, double* r,
                               float* r32, double* rr_out, double* partials);
    void launch_x_update_mixed(hipStream_t s, int64_t n, double* x, const float* d);
    // Vector ops (FP64); reductions write a device scalar.
    void launch_axpy(hipStream_t s, int64_t n, double* dst, const double* a, double sc, const double* b);
    void launch_scale(hipStream_t s, int64_t n, double* dst, const double* a, double sc);
    void launch_dot(hipStream_t s, int64_t n, const double* a, const double* b, double* out, double* partials);
    void launch_max_abs(hipStream_t s, int64_t n, const double* a, double* out, double* partials);
    void launch_fill_u8(hipStream_t s, int64_t n, uint8_t* p, uint8_t v);
    int reduction_groups(int64_t n);
    // Allocation-free loop evaluation: out[c * n_elem + e] = in[e * stride + c], and the device
    // address of the per-GPU scratch array (device must be current; 0 on success).
    void launch_transpose_conn(hipStream_t s, const int32_t* in, int32_t n_elem, int32_t stride, int32_t* out);
    int loop_scratch_address(int device, int32_t** out);
}
