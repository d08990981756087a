// Solver front end mirroring SymX's symx/src/solver/{GlobalPotential,NewtonsMethod,
// solver_utils,Context,fem_integrators,mesh_generation}.h. The control flow of
// NewtonsMethod::solve is the reference's (NewtonsMethod.cpp:28-641); every
// vector and matrix operation runs on the GPU through include/symx_hip.h.
#pragma once
#include <array>
#include <chrono>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/symx_hip.h"
#include "compile.h"
#include "symbol.h"

namespace symx
{
    // ---- logging (Logger.h subset: timers and per-solve series) ----------------------
    class Logger
    {
    public:
        struct Stats { double total = 0, avg = 0, min = 0, max = 0; int64_t count = 0; };
        class ScopedTimer
        {
        public:
            ScopedTimer(Logger* l, std::string label) : l(l), label(std::move(label)), t0(std::chrono::steady_clock::now()) {}
            ~ScopedTimer() { l->timers[label] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
        private:
            Logger* l;
            std::string label;
            std::chrono::steady_clock::time_point t0;
        };
        ScopedTimer time(const std::string& label) { return ScopedTimer(this, label); }
        void add_and_append(const std::string& key, double v) { series[key].push_back(v); }
        Stats get_stats(const std::string& key) const;
        std::vector<std::string> get_timer_labels() const;
        double get_timer_total(const std::string& label) const;
        void clear() { timers.clear(); series.clear(); }
        std::map<std::string, double> timers;
        std::map<std::string, std::vector<double>> series;
    };
    using spLogger = std::shared_ptr<Logger>;

    enum class Verbosity { None = 0, Summary = 1, Medium = 2, Full = 3 };
    class OutputSink
    {
    public:
        void print(const std::string& s, Verbosity v = Verbosity::Medium);
        void print_with_new_line(const std::string& s, Verbosity v = Verbosity::Medium);
        void print_new_line(Verbosity v = Verbosity::Medium) { print("\n", v); }
        void set_console_verbosity(Verbosity v) { verbosity = v; }
        Verbosity get_console_verbosity() const { return verbosity; }
        void set_enabled(bool on) { enabled = on; }
    private:
        Verbosity verbosity = Verbosity::None;
        bool enabled = true;
    };
    using spOutputSink = std::shared_ptr<OutputSink>;

    // ---- context (Context.h) -----------------------------------------------------------
    struct Context
    {
        int n_threads = 1;
        std::string compilation_directory;  // code-object cache
        int device = 0;
        spOutputSink output = std::make_shared<OutputSink>();
        spLogger logger = std::make_shared<Logger>();
        static std::shared_ptr<Context> create() { return std::make_shared<Context>(); }
        // Lazily created HIP context shared by everything built from this Context.
        symx_hip_ctx hip();
        ~Context();
    private:
        symx_hip_ctx hip_ctx = nullptr;
    };
    using spContext = std::shared_ptr<Context>;
    std::string get_codegen_dir();

    // ---- settings, callbacks, return codes (solver_utils.h:15-258) --------------------
    enum class SolverReturn
    {
        Successful,
        Running,
        InvalidInitialState,
        TooManyIterations,
        TooManyArmijoIterations,
        LinearSystemSolveFailure,
        TooManyInvalidIntermediateIterations,
        StepDoesNotDescend,
        InvalidConvergedState,
    };
    const char* to_string(SolverReturn r);

    enum class LinearSolver { DirectLLT, BDPCG };
    enum class ProjectionToPD { Newton, ProjectedNewton, ProjectOnDemand, Progressive };
    // Additive backend knob: FP64 everywhere, or FP32 element Hessians / block matrix /
    // PCG inside FP64 iterative refinement.
    enum class Precision { FP64, Mixed };

    struct SolverSettings
    {
        int max_iterations = std::numeric_limits<int>::max();
        int min_iterations = 0;
        double residual_tolerance_abs = 1e-6;
        double residual_tolerance_rel = 0.0;
        double step_tolerance = 0.0;
        bool max_iterations_as_success = false;
        double step_cap = std::numeric_limits<double>::infinity();
        bool enable_armijo_backtracking = true;
        double line_search_armijo_beta = 1e-4;
        int max_backtracking_armijo_iterations = 20;
        int max_backtracking_invalid_state_iterations = 8;
        bool print_line_search_upon_failure = false;
    };

    struct NewtonSettings : public SolverSettings
    {
        ProjectionToPD projection_mode = ProjectionToPD::ProjectedNewton;
        double projection_eps = 1e-10;
        bool project_to_pd_use_mirroring = false;
        int project_on_demand_countdown = 4;
        double ppn_tightening_factor = 0.5;
        double ppn_release_factor = 2.0;
        LinearSolver linear_solver = LinearSolver::BDPCG;
        int cg_max_iterations = 10000;
        double cg_abs_tolerance = 1e-12;
        double cg_rel_tolerance = 1e-4;
        bool cg_stop_on_indefiniteness = true;
        double bailout_residual = 1e-10;
        Precision precision = Precision::FP64;
    };

    // Residual callbacks receive the gradient downloaded to the host.
    class SolverCallbacks
    {
    public:
        explicit SolverCallbacks(spContext c) : context(c) {}
        static std::shared_ptr<SolverCallbacks> create(spContext c) { return std::make_shared<SolverCallbacks>(c); }
        void set_residual(std::function<double(std::vector<double>&)> f) { residual = f; has_custom_residual = true; }
        void add_before_energy_evaluation(std::function<void()> f) { before_energy_evaluation.push_back(f); }
        void add_is_initial_state_valid(std::function<bool()> f) { is_initial_state_valid.push_back(f); }
        void add_is_intermediate_state_valid(std::function<bool()> f) { is_intermediate_state_valid.push_back(f); }
        void add_on_intermediate_state_invalid(std::function<void()> f) { on_intermediate_state_invalid.push_back(f); }
        void add_on_armijo_fail(std::function<void()> f) { on_armijo_fail.push_back(f); }
        void add_is_converged(std::function<bool()> f) { is_converged.push_back(f); }
        void add_is_converged_state_valid(std::function<bool()> f) { is_converged_state_valid.push_back(f); }
        void add_max_allowed_step(std::function<double()> f) { max_allowed_step.push_back(f); }
        // Device-side variant (no counterpart in SymX): called with the device problem,
        // whose VEC_DOFS holds the DoFs before the step and VEC_DU the (capped) step, e.g.
        // symx_hip_problem_plane_max_step. No host synchronisation of the DoFs.
        void add_max_allowed_step_device(std::function<double(symx_hip_problem)> f) { max_allowed_step_device.push_back(f); }

        std::vector<std::function<void()>> before_energy_evaluation;
        std::vector<std::function<bool()>> is_initial_state_valid;
        std::vector<std::function<bool()>> is_intermediate_state_valid;
        std::vector<std::function<void()>> on_intermediate_state_invalid;
        std::vector<std::function<void()>> on_armijo_fail;
        std::vector<std::function<bool()>> is_converged;
        std::vector<std::function<bool()>> is_converged_state_valid;
        std::vector<std::function<double()>> max_allowed_step;
        std::vector<std::function<double(symx_hip_problem)>> max_allowed_step_device;
        std::function<double(std::vector<double>&)> residual;
        bool has_custom_residual = false;
        spContext context;
    };
    using spSolverCallbacks = std::shared_ptr<SolverCallbacks>;

    // ---- potentials (Potential.h, GlobalPotential.h) ------------------------------------
    class Potential
    {
    public:
        std::string name;
        spMWS<double> mws;
        Scalar expression;
        Scalar condition;
        bool has_condition = false;
        // Graph nodes of the adjoint symbols (one per DoF symbol), created on first use.
        mutable std::vector<int32_t> adjoint_symbol_nodes;
        // Graph nodes of the symbols standing for the linear cut of the expression (the
        // chain-rule factorisation of its derivatives), created on first use.
        mutable std::vector<int32_t> cut_symbol_nodes;
    };

    class GlobalPotential
    {
    public:
        using DefFunc = std::function<Scalar(MappedWorkspace<double>&, Element&)>;
        using DefWCondFunc = std::function<std::pair<Scalar, Scalar>(MappedWorkspace<double>&, Element&)>;

        static std::shared_ptr<GlobalPotential> create() { return std::make_shared<GlobalPotential>(); }

        void add_potential(const std::string& name, spMWS<double> mws, Element& elem, DefFunc f);
        void add_potential(const std::string& name, spMWS<double> mws, Element& elem, DefWCondFunc f);
        void add_potential(const std::string& name, std::function<const int32_t*()> data, std::function<int32_t()> n_elements,
                           int32_t stride, DefFunc f);
        void add_potential(const std::string& name, const std::vector<int32_t>& arr, int32_t n_items_per_element, DefFunc f);
        void add_potential(const std::string& name, const std::vector<int32_t>& arr, int32_t n_items_per_element, DefWCondFunc f);
        template <std::size_t N>
        void add_potential(const std::string& name, const std::vector<std::array<int32_t, N>>& arr, DefFunc f)
        {
            auto [mws, elem] = MappedWorkspace<double>::create(arr);
            add_potential(name, mws, elem, f);
        }
        template <std::size_t N>
        void add_potential(const std::string& name, const std::vector<std::array<int32_t, N>>& arr, DefWCondFunc f)
        {
            auto [mws, elem] = MappedWorkspace<double>::create(arr);
            add_potential(name, mws, elem, f);
        }
        template <std::size_t N>
        void add_potential(const std::string& name, const LabelledConnectivity<N>& conn, DefFunc f)
        {
            auto [mws, elem] = MappedWorkspace<double>::create(conn);
            add_potential(name, mws, elem, f);
        }
        template <std::size_t N>
        void add_potential(const std::string& name, const LabelledConnectivity<N>& conn, DefWCondFunc f)
        {
            auto [mws, elem] = MappedWorkspace<double>::create(conn);
            add_potential(name, mws, elem, f);
        }

        void add_dof(std::function<std::uintptr_t()> id, std::function<double*()> data, std::function<int32_t()> flat_size,
                     const std::string& name = "");
        template <typename STATIC_VECTOR>
        void add_dof(std::vector<STATIC_VECTOR>& arr, const std::string& name = "")
        {
            constexpr int32_t stride = (int32_t)(sizeof(STATIC_VECTOR) / sizeof(double));
            add_dof([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return reinterpret_cast<double*>(arr.data()); },
                    [&arr]() { return (int32_t)(arr.size() * stride); }, name);
        }
        // Any contiguous container of doubles with data() and size(): a dynamic vector, or a
        // row-major matrix (GlobalPotential.h:55-61 takes EigenMatrixRM<double> this way).
        template <typename DYNAMIC_VECTOR>
        void add_dof(DYNAMIC_VECTOR& arr, const std::string& name = "")
        {
            add_dof([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return reinterpret_cast<double*>(arr.data()); },
                    [&arr]() { return (int32_t)arr.size(); }, name);
        }
        void add_dof(double& scalar, const std::string& name = "")
        {
            add_dof([&scalar]() { return reinterpret_cast<std::uintptr_t>(&scalar); }, [&scalar]() { return &scalar; }, []() { return (int32_t)1; },
                    name);
        }

        // Parameters of an adjoint sensitivity analysis (no counterpart in SymX): data
        // containers, identified like DoF sets by their address, whose symbols become the
        // p of d^2E/dx dp. Potentials bind them with the ordinary make_scalar/make_vector.
        void add_parameter(std::function<std::uintptr_t()> id, std::function<double*()> data, std::function<int32_t()> flat_size,
                           const std::string& name = "");
        template <typename DYNAMIC_VECTOR>
        void add_parameter(DYNAMIC_VECTOR& arr, const std::string& name = "")
        {
            constexpr int32_t stride = (int32_t)(sizeof(typename DYNAMIC_VECTOR::value_type) / sizeof(double));
            add_parameter([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return reinterpret_cast<double*>(arr.data()); },
                          [&arr]() { return (int32_t)(arr.size() * stride); }, name);
        }
        const std::vector<int32_t>& get_parameter_offsets();
        int32_t get_total_n_parameters() const;

        const std::vector<int32_t>& get_dofs_offsets();
        int32_t get_n_dof_sets() const { return (int32_t)dof_maps.size(); }
        int32_t get_n_dofs(int32_t set_i) const { return dof_maps[set_i].flat_size(); }
        int32_t get_total_n_dofs() const;
        void get_dofs(double* u) const;
        void set_dofs(const double* u);
        void apply_dof_increment(const double* du);
        const std::vector<std::unique_ptr<Potential>>& get_potentials() const { return potentials; }
        int32_t get_n_potentials() const { return (int32_t)potentials.size(); }

        struct DofSet
        {
            std::function<std::uintptr_t()> id;
            std::function<double*()> data;
            std::function<int32_t()> flat_size;
            std::string name;
        };
        const std::vector<DofSet>& get_dof_sets() const { return dof_maps; }
        const std::vector<DofSet>& get_parameter_sets() const { return param_maps; }

    private:
        std::vector<std::unique_ptr<Potential>> potentials;
        std::vector<DofSet> dof_maps;
        std::vector<int32_t> dof_offsets;
        std::vector<DofSet> param_maps;
        std::vector<int32_t> param_offsets;
    };
    using spGlobalPotential = std::shared_ptr<GlobalPotential>;

    // ---- device-compiled global potential (SecondOrderCompiledGlobal) ------------------
    // Owns the generated modules and the device problem.
    class DeviceGlobal
    {
    public:
        DeviceGlobal(spGlobalPotential G, spContext ctx);
        ~DeviceGlobal();
        symx_hip_problem problem() { return prob; }
        // Re-upload non-DoF data maps from the user arrays (they may have changed).
        void refresh_data();
        // Whether the potentials' element counts / DoF counts still match the device problem.
        bool matches_topology() const;
        int32_t n_dofs() const { return ndofs; }
        // Where output j of a potential's adjoint kernel lands in dL/dp.
        struct ParamOut
        {
            int32_t set = 0;       // parameter set
            int32_t conn_col = -1; // connectivity column of the data map (-1: one value for all elements)
            int32_t stride = 1;    // values per item of the data map
            int32_t k = 0;         // component within the item
        };
        struct PotentialInfo
        {
            std::string name;
            int32_t n_elem = 0;
            int32_t n_dofs = 0;
            int64_t ops_P = 0, ops_PG = 0, ops_PGH = 0, ops_adjoint = 0;
            std::string source;
            std::string adjoint_kernel;          // empty: the potential depends on no parameter
            std::vector<ParamOut> param_outs;
        };
        struct AdjointStats
        {
            int cg_iterations = 0;
            int n_refinements = 0;
            bool converged = false;
            bool found_indefiniteness = false;
            double cg_error = 0.0;
            double energy = 0.0;
            double seconds = 0.0;
        };
        // Adjoint sensitivity at the DoFs currently on the device (an equilibrium):
        // H lambda = dL_dx with H the unprojected Hessian, then
        // dL_dp = -lambda^T d^2E/dx dp over all registered parameter sets (in their
        // order, flattened). The explicit dL/dp of the loss is not included.
        AdjointStats adjoint(const double* dL_dx, double* dL_dp, Precision precision, double rel_tol, int32_t max_iter);
        int32_t n_parameters() const { return nparams; }
        const std::vector<PotentialInfo>& info() const { return infos; }
        double compile_seconds = 0.0;
        int n_cached = 0;

    private:
        spGlobalPotential G;
        spContext ctx;
        symx_hip_problem prob = nullptr;
        std::vector<symx_hip_module> modules;
        std::vector<PotentialInfo> infos;
        int32_t ndofs = 0;
        int32_t nparams = 0;
        std::vector<int32_t> param_offsets;
        struct HostMap { int32_t pot, map; std::function<const double*()> data; int64_t n; };
        std::vector<HostMap> host_maps;
    };

    // Generated kernels of one potential (P, PG, PGH, optional C and adjoint) and the
    // DoF layout they assume. Needs no device.
    struct PotentialModule
    {
        std::vector<KernelSpec> specs;
        std::string source;
        DeviceGlobal::PotentialInfo info;
        std::vector<int32_t> dof_col;        // connectivity column per DoF block
        std::vector<int32_t> dof_block_set;  // DoF set per DoF block
        std::vector<int32_t> map_dof_set;    // DoF set per data map (-1: not a DoF)
        bool structured = false;             // structured FEM kernels (see FemStructure)
        bool factorised = false;             // derivatives through the linear cut (chain_rule_derivatives)
        int32_t h_stride = 0;                // values per element Hessian record (0: n_ub * 9)
        FemStructure fem;
        int32_t elems_per_group = 256;
    };
    bool build_fem_structure(GlobalPotential& G, const Potential& P, const std::vector<Scalar>& dofs, FemStructure& fem);
    // Structured FEM kernels for elements of 8+ nodes (on by default). Off: every
    // fem_integrator expression is expanded and compiled as a generic element kernel.
    void set_structured_fem(bool enabled);
    bool get_structured_fem();
    // Derivatives of generic potentials through the linear cut of their expression (chain
    // rule, used when it takes fewer operations than the flat derivatives); on by default.
    void set_chain_rule(bool enabled);
    bool get_chain_rule();
    // contraction_inputs: the nodes the Hessian contraction reads (second derivatives of the
    // cut quantities and the nonzero coefficients dL/dx).
    bool chain_rule_derivatives(Graph& g, const Potential& P, const Scalar& expression, const std::vector<Scalar>& dofs, DiffCache& dc,
                                Vector& grad, Matrix& hess, std::vector<int32_t>* contraction_inputs = nullptr);
    PotentialModule generate_potential_module(GlobalPotential& G, const Potential& P);

    // ---- views (View.h) ----------------------------------------------------------------
    template <typename T>
    class View
    {
    public:
        View(T* data, int32_t size) : d(data), n(size) {}
        int32_t size() const { return n; }
        T* data() { return d; }
        const T* data() const { return d; }
        T& operator[](int32_t i) { return d[i]; }
        const T& operator[](int32_t i) const { return d[i]; }
        View slice(int32_t begin, int32_t end) const { return View(d + begin, end - begin); }
        View slice(int32_t begin) const { return View(d + begin, n - begin); }

    private:
        T* d;
        int32_t n;
    };

    // ---- host copy of the assembled Hessian (BlockedSparseMatrix<3,3,double> subset) ------
    struct Triplet
    {
        int32_t r = 0, c = 0;
        double v = 0.0;
        int32_t row() const { return r; }
        int32_t col() const { return c; }
        double value() const { return v; }
    };
    class BlockedSparseMatrix
    {
    public:
        int32_t n_block_rows = 0;
        std::vector<int64_t> rows;  // block-row offsets (n_block_rows + 1)
        std::vector<int32_t> cols;  // block column of each 3x3 block
        std::vector<double> vals;   // 9 row-major values per block
        void to_triplets(std::vector<Triplet>& triplets) const;
        std::vector<double> to_dense() const;  // row-major (3n x 3n)
        void spmxv(const double* x, double* y) const;
    };
    using spBSM = std::shared_ptr<BlockedSparseMatrix>;

    // Element Hessians of the last evaluation, resident on the device (ElementHessians.h).
    class ElementHessians
    {
    public:
        ElementHessians(symx_hip_problem p, symx_hip_ctx c) : prob(p), ctx(c) {}
        // project_to_PD.cpp semantics on every element Hessian; returns the number changed
        int64_t project_to_PD_inplace__all(double projection_eps, bool use_mirroring);
        // FP64 assembly on the device, copied to the host (n_threads is unused)
        spBSM assemble_global(int n_threads, int ndofs);

    private:
        symx_hip_problem prob;
        symx_hip_ctx ctx;
    };
    using spElementHessians = std::shared_ptr<ElementHessians>;

    // SecondOrderCompiledGlobal.h: evaluation of a GlobalPotential at the DoFs and data
    // currently in the user's arrays (uploaded before each evaluation).
    class SecondOrderCompiledGlobal
    {
    public:
        SecondOrderCompiledGlobal(spGlobalPotential G, spContext context);
        void evaluate_P(double& out_P);
        void evaluate_P__dP_du(double& out_P, std::vector<double>& out_dP_du);
        spElementHessians evaluate_P__dP_du__local_d2P_du2(double& out_P, std::vector<double>& out_dP_du);
        // Central differences of E -> gradient and gradient -> Hessian against the
        // compiled derivatives; returns the larger max abs difference
        // (SecondOrderCompiledGlobal.cpp:144-191). Dense: small problems only.
        double test_derivatives_with_FD(double h);
        DeviceGlobal& device() { return *dev; }

    private:
        void upload();
        spGlobalPotential G;
        spContext context;
        std::unique_ptr<DeviceGlobal> dev;
    };

    // CompiledInLoop<double> (CompiledInLoop.h:100-135): per-element outputs of a
    // mapped workspace evaluated on the device, then handed to a host callback per
    // element in element order (thread_id 0; n_threads is unused).
    template <typename T>
    class CompiledInLoop;
    template <>
    class CompiledInLoop<double>
    {
    public:
        using Callback = std::function<void(View<double> solution, int32_t element_idx, int32_t thread_id, View<int32_t> element_conn)>;
        CompiledInLoop(const spMWS<double>& mws, const std::vector<Scalar>& expr, const std::string& name, const std::string& folder = "",
                       const std::string& cache_id = "");
        ~CompiledInLoop();
        static std::shared_ptr<CompiledInLoop<double>> create(const spMWS<double>& mws, const std::vector<Scalar>& expr, const std::string& name,
                                                              const std::string& folder = "", const std::string& cache_id = "")
        {
            return std::make_shared<CompiledInLoop<double>>(mws, expr, name, folder, cache_id);
        }
        void run(int32_t n_threads, const Callback& callback);
        void run(int32_t n_threads, const Callback& callback, const std::vector<uint8_t>& is_element_active);
        bool was_cached() const { return cached; }
        int32_t get_n_outputs() const { return n_out; }
        const std::string& source() const { return src; }

    private:
        spMWS<double> mws;
        std::string kernel, src;
        int32_t n_out = 0;
        bool cached = false;
        symx_hip_ctx ctx = nullptr;
        symx_hip_module mod = nullptr;
        std::vector<double> out;
    };

    class NewtonsMethod
    {
    public:
        struct SolveStats
        {
            int newton_iterations = 0;
            int cg_iterations = 0;
            int ls_cap_iterations = 0;
            int ls_max_iterations = 0;
            int ls_inv_iterations = 0;
            int ls_bt_iterations = 0;
            int ls_trapezoid_accepts = 0;   // line searches ended by the trapezoid estimate of the energy change
            uint64_t n_hessians = 0;
            uint64_t n_projected_hessians = 0;
            double projected_hessians_ratio = 0.0;
            int n_refinements = 0;
            double final_residual = 0.0;
            double final_energy = 0.0;
            std::vector<double> energies;   // E at each Newton iteration
            std::vector<double> residuals;  // residual at each Newton iteration
            std::vector<int> cg_per_iteration;
            std::vector<double> step_sizes;
        };

        NewtonsMethod(spGlobalPotential G, spContext context, spSolverCallbacks callbacks = nullptr);
        static std::shared_ptr<NewtonsMethod> create(spGlobalPotential G, spContext context, spSolverCallbacks callbacks = nullptr)
        {
            return std::make_shared<NewtonsMethod>(G, context, callbacks);
        }
        SolverReturn solve();
        // Adjoint gradient dL/dp at the current DoFs of the user arrays, normally the
        // equilibrium that solve() returned (see DeviceGlobal::adjoint). Solved in FP64
        // whatever settings.precision is.
        DeviceGlobal::AdjointStats adjoint(const std::vector<double>& dL_dx, std::vector<double>& dL_dp, double rel_tol = 1e-10,
                                           int32_t max_iter = 100000);
        const SolveStats& get_last_solve_stats() const { return stats; }
        void print_summary(double total_time = -1.0) const;
        DeviceGlobal& device() { return *dev; }

        spSolverCallbacks callbacks;
        NewtonSettings settings;

    private:
        bool project_and_assemble(bool& assembled, int& pdn_countdown, double& ppn_threshold);
        bool solve_linear_system(double residual_norm);
        SolverReturn line_search(double du_dot_grad, double du_max);
        void sync_dofs_to_host();
        void upload_dofs();
        double compute_residual();
        double energy();

        spGlobalPotential G;
        spContext context;
        std::unique_ptr<DeviceGlobal> dev;
        SolveStats stats;
        std::vector<double> host_buf;
        int last_cg_iterations = 0;
    };
    using spNewtonsMethod = std::shared_ptr<NewtonsMethod>;

    // ---- single-expression runtime (Compiled.h) ----------------------------------------
    // NaN checks of Compiled<double> inputs and outputs (static_methods.h:15-16; off by default).
    void check_mode_ON(bool check = true);
    bool get_check_mode_ON();

    // Compilation::Info (Compilation.h:47-57).
    struct CompilationInfo
    {
        int n_inputs = -1;
        int n_outputs = -1;
        bool was_cached = false;
        double runtime_codegen = 0.0;      // seconds to emit the code
        double runtime_compilation = 0.0;  // seconds in the host compiler
        int n_bytes_symbols = 0;
        void print() const;
    };

    // Compiled<double> (Compiled.h:44-91): a list of expressions of one workspace evaluated
    // for the values set by the caller. Single calls (run) execute host code compiled from
    // the expressions, like the reference's own JIT (a C++ translation unit built by the
    // system compiler, cached in `folder` and validated by its embedded SHA-256); batches
    // (run_batch) run on the GPU, whose module is built on the first batch. Per-thread input
    // and output buffers as in the reference (set_n_threads, the thread_id overloads).
    template <typename T>
    class Compiled;

    template <>
    class Compiled<double>
    {
    public:
        Compiled() = default;
        Compiled(const Compiled&) = delete;
        Compiled(const std::vector<Scalar>& expr, const std::string& name, const std::string& folder, const std::string& cache_id = "");
        ~Compiled();

        void compile(const std::vector<Scalar>& expr, const std::string& name, const std::string& folder, const std::string& cache_id = "");
        bool load_if_cached(const std::string& name, const std::string& folder, const std::string& cache_id);
        void try_load_otherwise_compile(const std::vector<Scalar>& expr, const std::string& name, const std::string& folder,
                                        const std::string& cache_id = "");

        bool is_valid() const { return host_f != nullptr; }
        bool was_cached() const { return info.was_cached; }
        CompilationInfo get_compilation_info() const { return info; }

        void set_n_threads(int n_threads);
        int get_n_threads() const { return n_threads; }

        void set(const Scalar& scalar, const double& v) { set(scalar, v, 0); }
        void set(const Scalar& scalar, const double& v, int thread_id);
        template <typename IndexableVector>
        void set(const Vector& vector, const IndexableVector& v)
        {
            set(vector, v, 0);
        }
        template <typename IndexableVector>
        void set(const Vector& vector, const IndexableVector& v, int thread_id)
        {
            for (int32_t i = 0; i < vector.size(); i++) set(vector[i], (double)v[i], thread_id);
        }
        void set(const Matrix& matrix, const double* v) { set(matrix, v, 0); }
        void set(const Matrix& matrix, const double* v, int thread_id);
        template <typename RowMajorParenthesisIndexableMatrix>
        typename std::enable_if<!std::is_pointer<RowMajorParenthesisIndexableMatrix>::value, void>::type set(const Matrix& matrix,
                                                                                                             const RowMajorParenthesisIndexableMatrix& m)
        {
            set(matrix, m, 0);
        }
        template <typename RowMajorParenthesisIndexableMatrix>
        typename std::enable_if<!std::is_pointer<RowMajorParenthesisIndexableMatrix>::value, void>::type
        set(const Matrix& matrix, const RowMajorParenthesisIndexableMatrix& m, int thread_id)
        {
            for (int32_t i = 0; i < matrix.rows(); i++)
                for (int32_t j = 0; j < matrix.cols(); j++) set(matrix(i, j), (double)m(i, j), thread_id);
        }

        View<double> run() { return run(0); }
        View<double> run(int thread_id);
        // Evaluate n input vectors on the GPU ([n][n_inputs] -> [n][n_outputs]).
        void run_batch(const double* inputs, int64_t n, double* outputs);

        int get_n_inputs() const { return n_in; }
        int get_n_outputs() const { return n_out; }
        std::string get_name() const { return name; }
        double* get_input_buffer(int thread_id = 0) { return thread_input.at(thread_id).data(); }
        double* get_output_buffer(int thread_id = 0) { return thread_output.at(thread_id).data(); }
        const double* get_input_buffer(int thread_id = 0) const { return thread_input.at(thread_id).data(); }
        const double* get_output_buffer(int thread_id = 0) const { return thread_output.at(thread_id).data(); }

        std::vector<std::vector<double>> thread_input;
        std::vector<std::vector<double>> thread_output;

    private:
        void release();
        void resize_buffers();
        bool load_files(const std::string& name, const std::string& folder, const std::string& sha);
        std::string name;
        std::string base;        // folder + "/" + name + "_" + first 16 digits of the SHA-256
        int n_threads = 1;
        int32_t n_in = 0, n_out = 0;
        CompilationInfo info;
        void* lib = nullptr;
        void (*host_f)(const double*, double*) = nullptr;
        std::string device_source;  // HIP source of the batch kernel (module built on the first batch)
        symx_hip_ctx ctx = nullptr;
        symx_hip_module mod = nullptr;
    };

    // ---- FEM helpers (fem_integrators.h) and meshes (mesh_generation.h) ----------------
    enum class FEM_Element { Tet4, Tet10, Hex8, Hex27 };
    std::string get_name(const FEM_Element& e);
    int get_stride(const FEM_Element& e);
    std::vector<std::array<double, 4>> get_integration_rule(const FEM_Element& e);
    Vector fem_interpolation(const FEM_Element& e, const std::vector<Vector>& xh, const Vector& xi);
    Matrix fem_jacobian(const FEM_Element& e, const std::vector<Vector>& xh, const Vector& xi);
    Scalar fem_integrator(MappedWorkspace<double>& mws, const FEM_Element& e, std::function<Scalar(Scalar& w, Vector& xi)> summand);

    struct FEMMesh
    {
        FEM_Element element_type = FEM_Element::Tet4;
        std::vector<std::array<double, 3>> vertices;
        std::vector<int32_t> connectivity;
        int32_t connectivity_stride = 4;
        int32_t n_elements = 0;
    };
    // Cuboid of `size` with its minimum corner at `origin`, split into Tet4 with the
    // 6-tetrahedra pattern of mesh_generation.cpp:78-83; vertices ordered i-major.
    FEMMesh generate_cuboid_Tet4_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& elements_per_axis,
                                      const std::array<double, 3>& origin = { 0.0, 0.0, 0.0 });
    // Cuboid of Tet10 elements: nodes on a (2n+1)^3 lattice, each cell split into the six
    // tetrahedra of the Tet4 mesh, element nodes v0..v3 then mid-edges 01 12 02 03 13 23
    // (mesh_generation.cpp:93-159; no RCM renumbering).
    FEMMesh generate_cuboid_Tet10_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& elements_per_axis,
                                       const std::array<double, 3>& origin = { 0.0, 0.0, 0.0 });
    // Cuboid of Hex8 elements (mesh_generation.cpp:161-206 node order; no RCM renumbering).
    FEMMesh generate_cuboid_Hex8_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& elements_per_axis,
                                      const std::array<double, 3>& origin = { 0.0, 0.0, 0.0 });
    // Cuboid of Hex27 elements (quadratic Lagrange, nodes on a (2n+1)^3 lattice, element node
    // order of mesh_generation.cpp:209-300; no RCM renumbering).
    FEMMesh generate_cuboid_Hex27_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& elements_per_axis,
                                       const std::array<double, 3>& origin = { 0.0, 0.0, 0.0 });
    // The generator of element type `element` (mesh_generation.cpp:307-326).
    FEMMesh generate_cuboid_mesh(const FEM_Element& element, const std::array<double, 3>& size, const std::array<int32_t, 3>& elements_per_axis,
                                 const std::array<double, 3>& origin = { 0.0, 0.0, 0.0 });
}
