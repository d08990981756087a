#include "solver.h"

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "../runtime/sha256.h"

namespace symx
{
    namespace
    {
        void check(int rc, symx_hip_ctx ctx, const char* what)
        {
            if (rc != SYMX_HIP_OK) throw std::runtime_error(std::string("symx_hip: ") + what + ": " + symx_hip_last_error(ctx));
        }
        std::string sfmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
        std::string sfmt(const char* f, ...)
        {
            char buf[256];
            va_list ap;
            va_start(ap, f);
            std::vsnprintf(buf, sizeof(buf), f, ap);
            va_end(ap);
            return buf;
        }
        // Line-search steps below this fraction of max|x| are judged by the trapezoid estimate
        // of the energy change when the energy difference does not show a decrease.
        constexpr double kTrapezoidStep = 1e-7;
        // Relative residual to which LinearSolver::DirectLLT solves (FP64 round-off level for the
        // penalty-stiff systems of the examples, condition numbers up to about 1e7).
        constexpr double kDirectSolveTol = 1e-12;
        std::string identifier(const std::string& s)
        {
            std::string r;
            for (char c : s) r += (std::isalnum((unsigned char)c) ? c : '_');
            if (r.empty() || std::isdigit((unsigned char)r[0])) r = "p_" + r;
            return r;
        }
    }

    // ---- Logger / OutputSink ------------------------------------------------------------
    Logger::Stats Logger::get_stats(const std::string& key) const
    {
        Stats s;
        auto it = series.find(key);
        if (it == series.end() || it->second.empty()) return s;
        s.min = it->second[0];
        s.max = it->second[0];
        for (double v : it->second) {
            s.total += v;
            s.min = std::min(s.min, v);
            s.max = std::max(s.max, v);
        }
        s.count = (int64_t)it->second.size();
        s.avg = s.total / (double)s.count;
        return s;
    }
    std::vector<std::string> Logger::get_timer_labels() const
    {
        std::vector<std::string> r;
        for (auto& kv : timers) r.push_back(kv.first);
        return r;
    }
    double Logger::get_timer_total(const std::string& label) const
    {
        auto it = timers.find(label);
        return it == timers.end() ? 0.0 : it->second;
    }
    void OutputSink::print(const std::string& s, Verbosity v)
    {
        if (enabled && (int)v <= (int)verbosity && verbosity != Verbosity::None) std::cout << s << std::flush;
    }
    void OutputSink::print_with_new_line(const std::string& s, Verbosity v) { print("\n" + s, v); }

    // ---- Context --------------------------------------------------------------------------
    symx_hip_ctx Context::hip()
    {
        if (!hip_ctx) check(symx_hip_ctx_create(device, &hip_ctx), nullptr, "creating the HIP context");
        return hip_ctx;
    }
    Context::~Context()
    {
        if (hip_ctx) symx_hip_ctx_destroy(hip_ctx);
    }
    std::string get_codegen_dir()
    {
        if (const char* d = std::getenv("SYMX_AMD_CODEGEN_DIR")) return d;
        if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/symx_amd";
        return "/tmp/symx_amd_codegen";
    }

    const char* to_string(SolverReturn r)
    {
        switch (r) {
        case SolverReturn::Successful: return "Successful";
        case SolverReturn::Running: return "Running";
        case SolverReturn::InvalidInitialState: return "InvalidInitialState";
        case SolverReturn::TooManyIterations: return "TooManyIterations";
        case SolverReturn::TooManyArmijoIterations: return "TooManyArmijoIterations";
        case SolverReturn::LinearSystemSolveFailure: return "LinearSystemSolveFailure";
        case SolverReturn::TooManyInvalidIntermediateIterations: return "TooManyInvalidIntermediateIterations";
        case SolverReturn::StepDoesNotDescend: return "StepDoesNotDescend";
        case SolverReturn::InvalidConvergedState: return "InvalidConvergedState";
        }
        return "?";
    }

    // ---- GlobalPotential (GlobalPotential.cpp) -------------------------------------------
    void GlobalPotential::add_potential(const std::string& name, spMWS<double> mws, Element& elem, DefFunc f)
    {
        auto p = std::make_unique<Potential>();
        p->name = name;
        p->mws = mws;
        p->expression = f(*mws, elem);
        potentials.push_back(std::move(p));
    }
    void GlobalPotential::add_potential(const std::string& name, spMWS<double> mws, Element& elem, DefWCondFunc f)
    {
        auto p = std::make_unique<Potential>();
        p->name = name;
        p->mws = mws;
        auto pc = f(*mws, elem);
        p->expression = pc.first;
        p->condition = pc.second;
        p->has_condition = true;
        potentials.push_back(std::move(p));
    }
    void GlobalPotential::add_potential(const std::string& name, std::function<const int32_t*()> data, std::function<int32_t()> n_elements,
                                        int32_t stride, DefFunc f)
    {
        auto [mws, elem] = MappedWorkspace<double>::create(ConnectivityMap(data, n_elements, stride));
        add_potential(name, mws, elem, f);
    }
    void GlobalPotential::add_potential(const std::string& name, const std::vector<int32_t>& arr, int32_t n, DefFunc f)
    {
        auto [mws, elem] = MappedWorkspace<double>::create(arr, n);
        add_potential(name, mws, elem, f);
    }
    void GlobalPotential::add_potential(const std::string& name, const std::vector<int32_t>& arr, int32_t n, DefWCondFunc f)
    {
        auto [mws, elem] = MappedWorkspace<double>::create(arr, n);
        add_potential(name, mws, elem, f);
    }
    void GlobalPotential::add_dof(std::function<std::uintptr_t()> id, std::function<double*()> data, std::function<int32_t()> flat_size,
                                  const std::string& name)
    {
        dof_maps.push_back(DofSet{ id, data, flat_size, name });
    }
    void GlobalPotential::add_parameter(std::function<std::uintptr_t()> id, std::function<double*()> data, std::function<int32_t()> flat_size,
                                        const std::string& name)
    {
        param_maps.push_back(DofSet{ id, data, flat_size, name });
    }
    const std::vector<int32_t>& GlobalPotential::get_parameter_offsets()
    {
        param_offsets.assign(param_maps.size() + 1, 0);
        for (size_t i = 0; i < param_maps.size(); i++) param_offsets[i + 1] = param_offsets[i] + param_maps[i].flat_size();
        return param_offsets;
    }
    int32_t GlobalPotential::get_total_n_parameters() const
    {
        int32_t n = 0;
        for (auto& d : param_maps) n += d.flat_size();
        return n;
    }
    const std::vector<int32_t>& GlobalPotential::get_dofs_offsets()
    {
        dof_offsets.assign(dof_maps.size() + 1, 0);
        for (size_t i = 0; i < dof_maps.size(); i++) dof_offsets[i + 1] = dof_offsets[i] + dof_maps[i].flat_size();
        return dof_offsets;
    }
    int32_t GlobalPotential::get_total_n_dofs() const
    {
        int32_t n = 0;
        for (auto& d : dof_maps) n += d.flat_size();
        return n;
    }
    void GlobalPotential::get_dofs(double* u) const
    {
        int32_t off = 0;
        for (auto& d : dof_maps) {
            const int32_t n = d.flat_size();
            std::copy(d.data(), d.data() + n, u + off);
            off += n;
        }
    }
    void GlobalPotential::set_dofs(const double* u)
    {
        int32_t off = 0;
        for (auto& d : dof_maps) {
            const int32_t n = d.flat_size();
            std::copy(u + off, u + off + n, d.data());
            off += n;
        }
    }
    void GlobalPotential::apply_dof_increment(const double* du)
    {
        int32_t off = 0;
        for (auto& d : dof_maps) {
            const int32_t n = d.flat_size();
            double* x = d.data();
            for (int32_t i = 0; i < n; i++) x[i] += du[off + i];
            off += n;
        }
    }

    // ---- structured FEM kernels ------------------------------------------------------------
    // The potential qualifies when it has no condition and no adjoint parameters, exactly
    // one deferred Jacobian is built from all of its DoFs (one DoF vector per node), every
    // other deferred Jacobian reads non-DoF data, each record's nodes come from one array
    // through the connectivity, and the DoFs enter the expression through that Jacobian only.
    bool build_fem_structure(GlobalPotential& G, const Potential& P, const std::vector<Scalar>& dofs, FemStructure& fem)
    {
        MappedWorkspace<double>& mws = *P.mws;
        Graph* g = mws.workspace().graph();
        const auto& maps = mws.get_maps();
        const SummationMap& sum = mws.get_summation();
        if (P.has_condition || sum.n_iterations <= 0 || sum.n_iterations > 64) return false;
        for (const auto& ps : G.get_parameter_sets())
            if (!mws.maps_of_container(ps.id()).empty()) return false;
        std::unordered_map<int32_t, int32_t> dof_pos;  // DoF symbol -> position in dofs
        for (int32_t i = 0; i < (int32_t)dofs.size(); i++) dof_pos[dofs[i].get_symbol_idx()] = i;
        // symbol -> (map, component)
        std::vector<std::array<int32_t, 2>> sym_map(g->n_symbols(), { -1, -1 });
        for (int32_t m = 0; m < (int32_t)maps.size(); m++)
            for (int k = 0; k < maps[m].stride; k++) sym_map[maps[m].first_symbol_idx + k] = { m, k };
        int dof_rec = -1;
        for (int r = 0; r < (int)mws.fem_jacobians.size(); r++) {
            const auto& rec = mws.fem_jacobians[r];
            int n_dof = 0;
            for (int32_t sidx : rec.node_symbols) n_dof += dof_pos.count(sidx) ? 1 : 0;
            if (n_dof == 0) continue;
            if (n_dof != (int)rec.node_symbols.size() || n_dof != (int)dofs.size() || dof_rec >= 0) return false;
            dof_rec = r;
        }
        if (dof_rec < 0) return false;
        // DoFs must not appear in the expression outside the deferred Jacobians
        {
            std::vector<uint8_t> seen(g->size(), 0);
            std::vector<int32_t> stack = { P.expression.id };
            while (!stack.empty()) {
                const int32_t id = stack.back();
                stack.pop_back();
                if (id < 0 || seen[id]) continue;
                seen[id] = 1;
                const Node& nd = g->node(id);
                if (nd.op == Op::Symbol) {
                    if (dof_pos.count(nd.a)) return false;
                    continue;
                }
                if (nd.op == Op::Zero || nd.op == Op::One || nd.op == Op::Const) continue;
                stack.push_back(nd.a);
                if (nd.op == Op::Branch) {
                    stack.push_back(nd.b);
                    stack.push_back(nd.c);
                }
                else if (nd.op != Op::PowN && nd.b >= 0) {
                    stack.push_back(nd.b);
                }
            }
        }
        const auto& drec = mws.fem_jacobians[dof_rec];
        const int NN = drec.n_nodes;
        std::vector<int> order = { dof_rec };
        for (int r = 0; r < (int)mws.fem_jacobians.size(); r++)
            if (r != dof_rec) order.push_back(r);
        fem = FemStructure();
        fem.n_nodes = NN;
        fem.n_q = sum.n_iterations;
        for (int r : order) {
            const auto& rec = mws.fem_jacobians[r];
            if (rec.n_nodes != NN) return false;
            FemStructure::Record fr;
            for (int k = 0; k < 9; k++) fr.J[k] = g->node(rec.J[k]).a;
            std::uintptr_t container = 0;
            for (int a = 0; a < NN; a++) {
                const auto mk = sym_map[rec.node_symbols[3 * a]];
                if (mk[0] < 0 || mk[1] != 0) return false;
                const DataMap& dm = maps[mk[0]];
                if (dm.stride != 3 || dm.connectivity_index < 0) return false;
                for (int k = 1; k < 3; k++)
                    if (sym_map[rec.node_symbols[3 * a + k]] != std::array<int32_t, 2>{ mk[0], k }) return false;
                if (a == 0) {
                    container = dm.id();
                    fr.map_base = mk[0];
                }
                else if (dm.id() != container) {
                    return false;
                }
                fr.cols.push_back(dm.connectivity_index);
            }
            fem.records.push_back(fr);
        }
        // DoF block of each node
        fem.block_of_node.resize(NN);
        std::vector<int32_t> node_of_block(NN, -1);
        for (int a = 0; a < NN; a++) {
            const int32_t pos = dof_pos.at(drec.node_symbols[3 * a]);
            if (pos % 3 != 0 || dof_pos.at(drec.node_symbols[3 * a + 1]) != pos + 1 || dof_pos.at(drec.node_symbols[3 * a + 2]) != pos + 2) return false;
            fem.block_of_node[a] = pos / 3;
            node_of_block[pos / 3] = a;
        }
        for (int A = 0; A < NN; A++)
            for (int B = A; B < NN; B++) fem.ub_nodes.push_back({ node_of_block[A], node_of_block[B] });
        // dN_a/dxi_k at every quadrature point: the coefficients of x_a in the expansion
        {
            std::vector<int32_t> ids;
            for (int a = 0; a < NN; a++)
                for (int k = 0; k < 3; k++)
                    ids.push_back(diff(Scalar(g, drec.expanded[k]), Scalar(g, g->symbol_node(drec.node_symbols[3 * a]))).id);
            fem.dN.resize((size_t)fem.n_q * NN * 3);
            for (int q = 0; q < fem.n_q; q++) {
                for (int k = 0; k < sum.stride; k++) g->set_symbol_value(sum.first_symbol_idx + k, sum.data[(size_t)q * sum.stride + k]);
                const std::vector<double> v = g->eval_many(ids);
                for (size_t i = 0; i < v.size(); i++) fem.dN[(size_t)q * NN * 3 + i] = v[i];
            }
        }
        // material law and its J-derivatives
        std::vector<Scalar> Jd;
        for (int k = 0; k < 9; k++) Jd.push_back(Scalar(g, drec.J[k]));
        DiffCache dc;
        fem.f = { P.expression };
        fem.P = gradient(P.expression, Jd, dc).values();
        const Matrix A = gradient(Vector(fem.P), Jd, true, dc);
        for (int r = 0; r < 9; r++)
            for (int c = r; c < 9; c++) fem.A.push_back(A(r, c));
        return true;
    }

    // ---- chain rule through the linear cut of an expression --------------------------------
    // Most energies reach their DoFs through a few quantities that are affine in the DoFs
    // with DoF-free coefficients: the deformation gradient F = Ds Dm^-1 of an elastic
    // element (9 quantities for 12 DoFs), a distance vector, a strain. With L those
    // quantities (the "cut": affine nodes that feed a non-affine operation) and
    // E(x) = f(L(x)),
    //     dE/dx = C^T df/dL,   d2E/dx2 = C^T (d2f/dL2) C,   C = dL/dx (DoF-free),
    // so the generated code evaluates the material law's m first and m(m+1)/2 second
    // derivatives once and contracts them with C's nonzeros, instead of differentiating the
    // flat expression twice with respect to every DoF. The derivatives are the same
    // functions (the rounding differs). The factorised form keeps fewer values live in the
    // generated kernel, so it is used unless it takes over 1.5 times the flat form's
    // operations.
    namespace
    {
        bool g_chain_rule = true;

        // Affine-cut nodes of `root` with respect to `dofs`; false when there is none worth
        // factorising through (the expression is affine, or the cut is not narrower than
        // the DoFs).
        bool find_linear_cut(const Graph& g, int32_t root, const std::vector<Scalar>& dofs, std::vector<int32_t>& cut)
        {
            std::vector<uint8_t> is_dof(g.n_symbols(), 0);
            for (const Scalar& d : dofs) is_dof[d.get_symbol_idx()] = 1;
            // degree: 0 DoF-free, 1 affine in the DoFs with DoF-free coefficients, 2 otherwise
            std::vector<uint8_t> deg(root + 1, 0), seen(root + 1, 0);
            std::vector<int32_t> stack = { root };
            while (!stack.empty()) {
                const int32_t id = stack.back();
                stack.pop_back();
                if (id < 0 || seen[id]) continue;
                seen[id] = 1;
                const Node& n = g.node(id);
                if (n.op == Op::Zero || n.op == Op::One || n.op == Op::Const || n.op == Op::Symbol) continue;
                stack.push_back(n.a);
                if (n.op == Op::Branch) {
                    stack.push_back(n.b);
                    stack.push_back(n.c);
                }
                else if (n.op != Op::PowN && n.b >= 0) {
                    stack.push_back(n.b);
                }
            }
            std::vector<uint8_t> feeds_nonaffine(root + 1, 0);
            for (int32_t id = 0; id <= root; id++) {
                if (!seen[id]) continue;
                const Node& n = g.node(id);
                uint8_t d = 0;
                switch (n.op) {
                case Op::Zero: case Op::One: case Op::Const: d = 0; break;
                case Op::Symbol: d = is_dof[n.a] ? 1 : 0; break;
                case Op::Add: case Op::Sub: d = std::max(deg[n.a], deg[n.b]); break;
                case Op::Mul: d = deg[n.a] == 0 ? deg[n.b] : (deg[n.b] == 0 ? deg[n.a] : 2); break;
                case Op::Branch: d = deg[n.c] == 0 ? std::max(deg[n.a], deg[n.b]) : 2; break;
                case Op::PowN: d = deg[n.a] == 0 ? 0 : 2; break;
                case Op::PowF: d = (deg[n.a] == 0 && deg[n.b] == 0) ? 0 : 2; break;
                default: d = deg[n.a] == 0 ? 0 : 2; break;  // one-operand functions
                }
                deg[id] = d;
                if (d == 2) {
                    feeds_nonaffine[n.a] = 1;
                    if (n.op == Op::Branch) {
                        feeds_nonaffine[n.b] = 1;
                        feeds_nonaffine[n.c] = 1;
                    }
                    else if (n.op != Op::PowN && n.b >= 0) {
                        feeds_nonaffine[n.b] = 1;
                    }
                }
            }
            if (deg[root] != 2) return false;
            cut.clear();
            for (int32_t id = 0; id <= root; id++)
                if (seen[id] && deg[id] == 1 && feeds_nonaffine[id]) cut.push_back(id);
            return !cut.empty() && cut.size() < dofs.size();
        }
    }

    void set_chain_rule(bool enabled) { g_chain_rule = enabled; }
    bool get_chain_rule() { return g_chain_rule; }

    bool chain_rule_derivatives(Graph& g, const Potential& P, const Scalar& expression, const std::vector<Scalar>& dofs, DiffCache& dc,
                                Vector& grad, Matrix& hess, std::vector<int32_t>* contraction_inputs)
    {
        std::vector<int32_t> cut;
        if (!find_linear_cut(g, expression.id, dofs, cut)) return false;
        const int32_t m = (int32_t)cut.size(), n = (int32_t)dofs.size();
        while ((int32_t)P.cut_symbol_nodes.size() < m) P.cut_symbol_nodes.push_back(g.symbol());
        std::vector<Scalar> Y;
        std::unordered_map<int32_t, int32_t> to_y, to_cut;
        for (int32_t r = 0; r < m; r++) {
            Y.push_back(Scalar(&g, P.cut_symbol_nodes[r]));
            to_y[cut[r]] = P.cut_symbol_nodes[r];
            to_cut[P.cut_symbol_nodes[r]] = cut[r];
        }
        const Scalar f(&g, g.substitute({ expression.id }, to_y)[0]);
        // f must reach the DoFs through the cut only
        {
            std::vector<uint8_t> dof_sym(g.n_symbols(), 0);
            for (const Scalar& d : dofs) dof_sym[d.get_symbol_idx()] = 1;
            std::vector<uint8_t> seen(g.size(), 0);
            std::vector<int32_t> stack = { f.id };
            while (!stack.empty()) {
                const int32_t id = stack.back();
                stack.pop_back();
                if (id < 0 || seen[id]) continue;
                seen[id] = 1;
                const Node& nd = g.node(id);
                if (nd.op == Op::Symbol) {
                    if (dof_sym[nd.a]) return false;
                    continue;
                }
                if (nd.op == Op::Zero || nd.op == Op::One || nd.op == Op::Const) continue;
                stack.push_back(nd.a);
                if (nd.op == Op::Branch) {
                    stack.push_back(nd.b);
                    stack.push_back(nd.c);
                }
                else if (nd.op != Op::PowN && nd.b >= 0) {
                    stack.push_back(nd.b);
                }
            }
        }
        DiffCache dcy;
        const Vector gy = gradient(f, Y, dcy);
        const Matrix Ay = gradient(gy, Y, true, dcy);
        // back to the cut nodes: df/dL and d2f/dL2 as functions of the element's data
        std::vector<int32_t> roots;
        for (int32_t r = 0; r < m; r++) roots.push_back(gy[r].id);
        for (int32_t r = 0; r < m; r++)
            for (int32_t s = r; s < m; s++) roots.push_back(Ay(r, s).id);
        const std::vector<int32_t> back = g.substitute(roots, to_cut);
        std::vector<Scalar> gL(m);
        std::vector<std::vector<Scalar>> AL(m, std::vector<Scalar>(m));
        {
            size_t k = 0;
            for (int32_t r = 0; r < m; r++) gL[r] = Scalar(&g, back[k++]);
            for (int32_t r = 0; r < m; r++)
                for (int32_t s = r; s < m; s++) AL[r][s] = AL[s][r] = Scalar(&g, back[k++]);
        }
        // C = dL/dx (DoF-free) and its nonzeros per DoF
        std::vector<std::vector<std::pair<int32_t, Scalar>>> col(n);
        for (int32_t r = 0; r < m; r++) {
            for (int32_t d = 0; d < n; d++) {
                const Scalar c = diff(Scalar(&g, cut[r]), dofs[d], dc);
                if (!c.is_zero()) col[d].push_back({ r, c });
            }
        }
        std::vector<Scalar> gv(n);
        for (int32_t d = 0; d < n; d++) {
            Scalar acc(&g, 0);
            for (const auto& rc : col[d]) acc = acc + rc.second * gL[rc.first];
            gv[d] = acc;
        }
        // Hessian entries in the order of the upper 3x3-block storage (block rows of a
        // 3-DoF-per-node layout), so the generated code finishes them chunk by chunk
        hess = Matrix::zero({ n, n }, Scalar(&g, 0));
        std::vector<uint8_t> done((size_t)n * n, 0);
        auto entry = [&](int32_t d, int32_t e) {
            if (done[(size_t)d * n + e]) return;
            Scalar acc(&g, 0);
            for (const auto& rc : col[d])
                for (const auto& sc : col[e]) acc = acc + (rc.second * sc.second) * AL[rc.first][sc.first];
            hess(d, e) = acc;
            hess(e, d) = acc;
            done[(size_t)d * n + e] = done[(size_t)e * n + d] = 1;
        };
        if (n % 3 == 0) {
            const int32_t nb = n / 3;
            for (int32_t a = 0; a < nb; a++)
                for (int32_t b = a; b < nb; b++)
                    for (int32_t i = 0; i < 3; i++)
                        for (int32_t j = 0; j < 3; j++) entry(3 * a + i, 3 * b + j);
        }
        for (int32_t d = 0; d < n; d++)
            for (int32_t e = d; e < n; e++) entry(d, e);
        grad = Vector(gv);
        if (contraction_inputs) {
            contraction_inputs->clear();
            for (int32_t r = 0; r < m; r++)
                for (int32_t s = r; s < m; s++) contraction_inputs->push_back(AL[r][s].id);
            for (int32_t d = 0; d < n; d++)
                for (const auto& rc : col[d]) contraction_inputs->push_back(rc.second.id);
        }
        return true;
    }

    // ---- per-potential code generation (SecondOrderCompiledPotential.cpp:8-70) ----------
    PotentialModule generate_potential_module(GlobalPotential& G, const Potential& P)
    {
        PotentialModule pm;
        MappedWorkspace<double>& mws = *P.mws;
        Graph* g = mws.workspace().graph();
        const auto& maps = mws.get_maps();
        // DoF identification: maps whose container is a DoF set, in set order then
        // map order (SecondOrderCompiledPotential.cpp:10-33).
        std::vector<Scalar> dofs;
        pm.map_dof_set.assign(maps.size(), -1);
        std::vector<std::array<int32_t, 2>> dof_maps;  // {map, component}
        for (int32_t s = 0; s < G.get_n_dof_sets(); s++) {
            const std::uintptr_t id = G.get_dof_sets()[s].id();
            for (int32_t m : mws.maps_of_container(id)) {
                const DataMap& dm = maps[m];
                if (dm.stride != 3)
                    throw std::runtime_error("symx: expected DoF stride of 3, got " + std::to_string(dm.stride) + " in potential " + P.name);
                if (dm.connectivity_index < 0) throw std::runtime_error("symx: DoF data must be indexed by the connectivity (" + P.name + ")");
                for (int k = 0; k < 3; k++) {
                    dofs.push_back(Scalar(g, g->symbol_node(dm.first_symbol_idx + k)));
                    dof_maps.push_back({ m, k });
                }
                pm.dof_col.push_back(dm.connectivity_index);
                pm.dof_block_set.push_back(s);
                pm.map_dof_set[m] = s;
            }
        }
        if (dofs.empty()) throw std::runtime_error("symx: potential with zero DoFs: " + P.name);
        const int32_t n = (int32_t)dofs.size();
        const std::string base = identifier(P.name);

        // Deferred isoparametric Jacobians: structured kernels, or substitute them back.
        Scalar expression = P.expression;
        Scalar condition = P.condition;
        if (!mws.fem_jacobians.empty()) {
            if (build_fem_structure(G, P, dofs, pm.fem)) {
                pm.structured = true;
                DeviceGlobal::PotentialInfo& info = pm.info;
                info.name = P.name;
                info.n_dofs = n;
                pm.specs.resize(3);
                pm.specs[0] = KernelSpec{ base + "_P", KernelKind::P, pm.fem.f, n };
                pm.specs[1] = KernelSpec{ base + "_P__dP_du", KernelKind::PG, collect_scalars(std::vector<std::vector<Scalar>>{ pm.fem.f, pm.fem.P }), n };
                pm.specs[2] = KernelSpec{ base + "_P__dP_du__d2P_du2", KernelKind::PGH, collect_scalars(std::vector<std::vector<Scalar>>{ pm.fem.f, pm.fem.P, pm.fem.A }), n };
                pm.source = emit_hip_fem_module(base, *g, maps, mws.get_summation(), pm.fem, pm.specs[0].name, pm.specs[1].name, pm.specs[2].name,
                                                symx_hip_eval_prelude());
                info.ops_P = count_operations(*g, pm.specs[0].outputs);
                info.ops_PG = count_operations(*g, pm.specs[1].outputs);
                info.ops_PGH = count_operations(*g, pm.specs[2].outputs);
                info.source = pm.source;
                pm.elems_per_group = 4;
                return pm;
            }
            std::unordered_map<int32_t, int32_t> rep;
            for (const auto& rec : mws.fem_jacobians)
                for (int k = 0; k < 9; k++) rep[rec.J[k]] = rec.expanded[k];
            const std::vector<int32_t> r = g->substitute({ P.expression.id, P.has_condition ? P.condition.id : -1 }, rep);
            expression = Scalar(g, r[0]);
            if (P.has_condition) condition = Scalar(g, r[1]);
        }

        DiffCache dc;
        Vector grad = gradient(expression, dofs, dc);
        Matrix hess = gradient(grad, dofs, true, dc);
        std::vector<int32_t> contraction_inputs;
        {
            Vector grad_c;
            Matrix hess_c;
            if (g_chain_rule && chain_rule_derivatives(*g, P, expression, dofs, dc, grad_c, hess_c, &contraction_inputs)) {
                const int64_t flat = count_operations(*g, collect_scalars({ { expression }, grad.values(), hess.values() }));
                const int64_t fact = count_operations(*g, collect_scalars({ { expression }, grad_c.values(), hess_c.values() }));
                // the factorised form keeps fewer values live (the cut's derivatives and the
                // contraction coefficients instead of per-DoF chains): worth some more operations
                if (2 * fact <= 3 * flat) {
                    grad = grad_c;
                    hess = hess_c;
                    pm.factorised = true;
                }
                else {
                    contraction_inputs.clear();
                }
            }
        }
        std::vector<KernelSpec>& specs = pm.specs;
        specs.resize(3);
        specs[0] = KernelSpec{ base + "_P", KernelKind::P, { expression }, n };
        specs[1] = KernelSpec{ base + "_P__dP_du", KernelKind::PG, collect_scalars({ { expression }, grad.values() }), n };
        specs[2] = KernelSpec{ base + "_P__dP_du__d2P_du2", KernelKind::PGH, collect_scalars({ { expression }, grad.values(), hess.values() }), n };
        specs[2].f32_from = contraction_inputs;
        if (P.has_condition) specs.push_back(KernelSpec{ base + "_C", KernelKind::C, { condition }, n });

        DeviceGlobal::PotentialInfo& info = pm.info;
        info.name = P.name;
        info.n_dofs = n;

        // Adjoint kernel: the mixed partials lambda^T d^2E/dx dp, generated as
        // d(lambda . g)/dp with one adjoint symbol per DoF symbol, read where that DoF is read.
        std::vector<Scalar> params;
        for (int32_t s = 0; s < (int32_t)G.get_parameter_sets().size(); s++) {
            const std::uintptr_t id = G.get_parameter_sets()[s].id();
            for (int32_t m : mws.maps_of_container(id)) {
                const DataMap& dm = maps[m];
                for (int k = 0; k < dm.stride; k++) {
                    params.push_back(Scalar(g, g->symbol_node(dm.first_symbol_idx + k)));
                    info.param_outs.push_back(DeviceGlobal::ParamOut{ s, dm.connectivity_index, dm.stride, k });
                }
            }
        }
        if (!params.empty()) {
            KernelSpec ks;
            ks.name = base + "_adjoint";
            ks.kind = KernelKind::PGP;
            ks.n_dofs = n;
            // The adjoint symbols are created once per potential, so that regenerating the
            // module gives the same source (and code-object cache key).
            while ((int32_t)P.adjoint_symbol_nodes.size() < n) P.adjoint_symbol_nodes.push_back(g->symbol());
            Scalar lg = Scalar(g, g->constant(0.0));
            for (int32_t i = 0; i < n; i++) {
                const int32_t lid = P.adjoint_symbol_nodes[i];
                ks.adjoint_symbols.push_back({ g->node(lid).a, dof_maps[i][0], dof_maps[i][1] });
                lg = lg + Scalar(g, lid) * grad[i];
            }
            ks.outputs = gradient(lg, params, dc).values();
            info.adjoint_kernel = ks.name;
            info.ops_adjoint = count_operations(*g, ks.outputs);
            specs.push_back(ks);
        }
        // single-point PG/PGH kernels store their outputs chunk by chunk (emit_chunked_kernel)
        if (mws.get_summation().n_iterations <= 1) pm.h_stride = hessian_record_stride((n / 3) * (n / 3 + 1) / 2 * 9);
        pm.source = emit_hip_module(base, *g, maps, mws.get_summation(), specs, symx_hip_eval_prelude());
        info.ops_P = count_operations(*g, specs[0].outputs);
        info.ops_PG = count_operations(*g, specs[1].outputs);
        info.ops_PGH = count_operations(*g, specs[2].outputs);
        info.source = pm.source;
        return pm;
    }

    // ---- DeviceGlobal (SecondOrderCompiledGlobal + SecondOrderCompiledPotential) ---------
    DeviceGlobal::DeviceGlobal(spGlobalPotential G_, spContext ctx_) : G(G_), ctx(ctx_)
    {
        symx_hip_ctx hip = ctx->hip();
        ndofs = G->get_total_n_dofs();
        if (ndofs <= 0) throw std::runtime_error("symx: no degrees of freedom");
        if (ndofs % 3 != 0) throw std::runtime_error("symx: ndofs must be divisible by 3");
        const std::vector<int32_t> offsets = G->get_dofs_offsets();
        check(symx_hip_problem_create(hip, ndofs, &prob), hip, "problem_create");
        const std::string cache = ctx->compilation_directory.empty() ? get_codegen_dir() : ctx->compilation_directory;
        const auto t0 = std::chrono::steady_clock::now();
        int32_t pot_index = 0;
        for (const auto& pp : G->get_potentials()) {
            const Potential& P = *pp;
            MappedWorkspace<double>& mws = *P.mws;
            const auto& maps = mws.get_maps();
            PotentialModule pm = generate_potential_module(*G, P);
            PotentialInfo& info = pm.info;
            const std::vector<KernelSpec>& specs = pm.specs;
            const std::string& src = pm.source;
            const std::vector<int32_t>& map_dof_set = pm.map_dof_set;
            const std::vector<int32_t>& dof_col = pm.dof_col;
            std::vector<int32_t> dof_off;
            for (int32_t s : pm.dof_block_set) dof_off.push_back(offsets[s]);
            const int32_t n = info.n_dofs;

            symx_hip_module mod = nullptr;
            char key[80];
            int was_cached = 0;
            size_t code_size = 0;
            check(symx_hip_jit_compile(src.c_str(), "gfx950", cache.c_str(), key, &was_cached, &code_size), nullptr, ("JIT of " + P.name).c_str());
            n_cached += was_cached;
            check(symx_hip_module_load(hip, src.c_str(), cache.c_str(), &mod), hip, ("loading module of " + P.name).c_str());
            modules.push_back(mod);

            const int32_t n_elem = mws.conn.n_elements();
            info.n_elem = n_elem;
            std::vector<symx_hip_map_desc> mdesc(maps.size());
            for (size_t m = 0; m < maps.size(); m++) {
                const DataMap& dm = maps[m];
                if (map_dof_set[m] >= 0) {
                    mdesc[m].is_dof = 1;
                    mdesc[m].dof_offset = offsets[map_dof_set[m]];
                    mdesc[m].host = nullptr;
                    mdesc[m].n_values = 0;
                }
                else {
                    mdesc[m].is_dof = 0;
                    mdesc[m].dof_offset = 0;
                    mdesc[m].host = dm.data();
                    mdesc[m].n_values = (int64_t)dm.n_elements() * dm.stride;
                    host_maps.push_back(HostMap{ pot_index, (int32_t)m, dm.data, mdesc[m].n_values });
                }
            }
            symx_hip_potential_desc d{};
            d.module = mod;
            d.kernel_P = specs[0].name.c_str();
            d.kernel_PG = specs[1].name.c_str();
            d.kernel_PGH = specs[2].name.c_str();
            d.kernel_C = P.has_condition ? specs[3].name.c_str() : nullptr;
            const std::string k_pgh32 = specs[2].name + "_f32";  // emitted beside every PGH kernel
            d.kernel_PGH_f32 = k_pgh32.c_str();
            d.n_elem = n_elem;
            d.conn_stride = mws.conn.stride;
            d.conn = mws.conn.data();
            d.n_dof_blocks = n / 3;
            d.dof_conn_col = dof_col.data();
            d.dof_set_offset = dof_off.data();
            d.n_maps = (int32_t)mdesc.size();
            d.maps = mdesc.data();
            d.elems_per_group = pm.elems_per_group;
            d.h_stride = pm.h_stride;
            int32_t pid = -1;
            check(symx_hip_problem_add_potential(prob, &d, &pid), hip, ("adding potential " + P.name).c_str());
            infos.push_back(info);
            pot_index++;
        }
        param_offsets = G->get_parameter_offsets();
        nparams = param_offsets.back();
        check(symx_hip_problem_finalize(prob), hip, "problem_finalize");
        compile_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    DeviceGlobal::~DeviceGlobal()
    {
        if (prob) symx_hip_problem_destroy(prob);
        for (auto m : modules) symx_hip_module_unload(m);
    }

    void DeviceGlobal::refresh_data()
    {
        for (auto& hm : host_maps) check(symx_hip_problem_set_map_data(prob, hm.pot, hm.map, hm.data(), hm.n), ctx->hip(), "refreshing data");
    }

    DeviceGlobal::AdjointStats DeviceGlobal::adjoint(const double* dL_dx, double* dL_dp, Precision precision, double rel_tol, int32_t max_iter)
    {
        symx_hip_ctx hip = ctx->hip();
        const auto t0 = std::chrono::steady_clock::now();
        AdjointStats st;
        check(symx_hip_problem_eval(prob, SYMX_HIP_EVAL_PGH, &st.energy), hip, "adjoint: Hessian evaluation");
        symx_hip_pcg_info info{};
        const int32_t prec = precision == Precision::Mixed ? SYMX_HIP_MIXED : SYMX_HIP_FP64;
        check(symx_hip_problem_adjoint_solve(prob, prec, rel_tol, max_iter, dL_dx, nullptr, &info), hip, "adjoint: linear solve");
        st.cg_iterations = info.n_iterations;
        st.n_refinements = info.n_refinements;
        st.converged = info.converged != 0;
        st.found_indefiniteness = info.found_indefiniteness != 0;
        st.cg_error = info.error;
        std::fill(dL_dp, dL_dp + nparams, 0.0);
        const auto& pots = G->get_potentials();
        std::vector<double> out;
        for (int32_t pi = 0; pi < (int32_t)infos.size(); pi++) {
            const PotentialInfo& inf = infos[pi];
            if (inf.adjoint_kernel.empty() || inf.n_elem == 0) continue;
            const int32_t n_out = (int32_t)inf.param_outs.size();
            out.assign((size_t)inf.n_elem * n_out, 0.0);
            check(symx_hip_problem_launch_adjoint(prob, pi, inf.adjoint_kernel.c_str(), n_out, out.data()), hip, "adjoint kernel");
            const ConnectivityMap& conn = pots[pi]->mws->conn;
            const int32_t* c = conn.data();
            for (int32_t e = 0; e < inf.n_elem; e++) {
                for (int32_t j = 0; j < n_out; j++) {
                    const ParamOut& po = inf.param_outs[j];
                    const int64_t item = po.conn_col < 0 ? 0 : c[(int64_t)e * conn.stride + po.conn_col];
                    const int64_t idx = param_offsets[po.set] + item * po.stride + po.k;
                    if (idx < param_offsets[po.set] || idx >= param_offsets[po.set + 1])
                        throw std::runtime_error("symx: adjoint parameter index out of range in potential " + inf.name);
                    dL_dp[idx] -= out[(size_t)e * n_out + j];
                }
            }
        }
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return st;
    }

    bool DeviceGlobal::matches_topology() const
    {
        if (G->get_total_n_dofs() != ndofs) return false;
        const auto& pots = G->get_potentials();
        if (pots.size() != infos.size()) return false;
        for (size_t i = 0; i < pots.size(); i++)
            if (pots[i]->mws->conn.n_elements() != infos[i].n_elem) return false;
        return true;
    }

    // ---- NewtonsMethod (NewtonsMethod.cpp:12-641) ----------------------------------------
    NewtonsMethod::NewtonsMethod(spGlobalPotential G_, spContext context_, spSolverCallbacks cb) : callbacks(cb), G(G_), context(context_)
    {
        if (!callbacks) callbacks = std::make_shared<SolverCallbacks>(context);
        dev = std::make_unique<DeviceGlobal>(G, context);
    }

    void NewtonsMethod::sync_dofs_to_host()
    {
        // one DoF array (the usual case): copy straight into it, without a host staging copy
        const auto& sets = G->get_dof_sets();
        if (sets.size() == 1 && sets[0].flat_size() == dev->n_dofs()) {
            check(symx_hip_problem_get_dofs(dev->problem(), sets[0].data()), context->hip(), "get_dofs");
            return;
        }
        host_buf.resize(dev->n_dofs());
        check(symx_hip_problem_get_dofs(dev->problem(), host_buf.data()), context->hip(), "get_dofs");
        G->set_dofs(host_buf.data());
    }

    void NewtonsMethod::upload_dofs()
    {
        const auto& sets = G->get_dof_sets();
        if (sets.size() == 1 && sets[0].flat_size() == dev->n_dofs()) {
            check(symx_hip_problem_set_dofs(dev->problem(), sets[0].data()), context->hip(), "set_dofs");
            return;
        }
        host_buf.resize(dev->n_dofs());
        G->get_dofs(host_buf.data());
        check(symx_hip_problem_set_dofs(dev->problem(), host_buf.data()), context->hip(), "set_dofs");
    }

    double NewtonsMethod::compute_residual()
    {
        if (callbacks->has_custom_residual) {
            std::vector<double> g(dev->n_dofs());
            check(symx_hip_problem_download(dev->problem(), SYMX_HIP_VEC_GRAD, g.data()), context->hip(), "download gradient");
            return callbacks->residual(g);
        }
        double r = 0.0;
        check(symx_hip_problem_max_abs(dev->problem(), SYMX_HIP_VEC_GRAD, &r), context->hip(), "max_abs");
        return r;
    }

    DeviceGlobal::AdjointStats NewtonsMethod::adjoint(const std::vector<double>& dL_dx, std::vector<double>& dL_dp, double rel_tol, int32_t max_iter)
    {
        const int ndofs = G->get_total_n_dofs();
        if ((int)dL_dx.size() != ndofs) throw std::runtime_error("symx: adjoint right-hand side must have one entry per DoF");
        if (!dev->matches_topology()) dev = std::make_unique<DeviceGlobal>(G, context);
        if (dev->n_parameters() == 0) throw std::runtime_error("symx: adjoint needs parameters (GlobalPotential::add_parameter)");
        symx_hip_ctx hip = context->hip();
        dev->refresh_data();
        host_buf.resize(ndofs);
        G->get_dofs(host_buf.data());
        check(symx_hip_problem_set_dofs(dev->problem(), host_buf.data()), hip, "set_dofs");
        dL_dp.assign(dev->n_parameters(), 0.0);
        // Always FP64: with FP32 matrix storage the refined solution is that of the rounded
        // matrix, whose error (about 1e-4 relative on penalty-stiff systems) is far above
        // what a gradient check needs, and the FP32 solve takes more iterations.
        return dev->adjoint(dL_dx.data(), dL_dp.data(), Precision::FP64, rel_tol, max_iter);
    }

    SolverReturn NewtonsMethod::solve()
    {
        const int ndofs = G->get_total_n_dofs();
        if (ndofs <= 0) throw std::runtime_error("symx error NewtonsMethod::solve(): No degrees of freedom.");
        if (ndofs % 3 != 0) throw std::runtime_error("symx error NewtonsMethod::solve(): ndofs must be divisible by 3.");
        if (!dev->matches_topology()) dev = std::make_unique<DeviceGlobal>(G, context);
        symx_hip_ctx hip = context->hip();
        symx_hip_problem prob = dev->problem();
        auto& logger = context->logger;
        auto& out = context->output;

        dev->refresh_data();
        upload_dofs();

        stats = SolveStats();
        double E0 = 0.0;
        double du_dot_grad = 0.0;
        double res_0 = std::numeric_limits<double>::max();
        SolverReturn result = SolverReturn::Running;
        int pdn_countdown = 0;
        double ppn_threshold = -1.0;
        const bool precision_mixed = settings.precision == Precision::Mixed;
        int64_t n_hessians = 0;
        symx_hip_problem_n_element_hessians(prob, &n_hessians);

        auto run_bool = [](bool def, const std::vector<std::function<bool()>>& fs) {
            bool v = def;
            for (auto& f : fs) v = v && f();
            return v;
        };

        if (!run_bool(true, callbacks->is_initial_state_valid)) {
            out->print_with_new_line("Newton failure: Invalid initial state.", Verbosity::Medium);
            result = SolverReturn::InvalidInitialState;
        }

        int newton_iteration = -1;
        const bool full = out->get_console_verbosity() == Verbosity::Full;
        while (result == SolverReturn::Running) {
            newton_iteration++;
            if (newton_iteration == settings.max_iterations) {
                out->print_with_new_line("Newton failure: Too many iterations.", Verbosity::Medium);
                result = settings.max_iterations_as_success ? SolverReturn::Successful : SolverReturn::TooManyIterations;
                break;
            }
            out->print_with_new_line(sfmt("%2d. ", newton_iteration), Verbosity::Medium);
            if (!callbacks->before_energy_evaluation.empty()) {
                sync_dofs_to_host();
                for (auto& f : callbacks->before_energy_evaluation) f();
                dev->refresh_data();
            }
            {
                auto t = logger->time("evaluate_P__dP_du__local_d2P_du2");
                // mixed precision stores the element Hessians as floats (FP64 arithmetic)
                check(symx_hip_problem_eval(prob, precision_mixed ? (int32_t)SYMX_HIP_EVAL_PGH_F32 : (int32_t)SYMX_HIP_EVAL_PGH, &E0), hip, "evaluation");
            }
            const double residual_norm = compute_residual();
            out->print(sfmt("r0: %.2e | ", residual_norm), Verbosity::Medium);
            stats.energies.push_back(E0);
            stats.residuals.push_back(residual_norm);
            if (newton_iteration == 0) res_0 = residual_norm;
            if (residual_norm < settings.bailout_residual) {
                result = SolverReturn::Successful;
                break;
            }
            if (newton_iteration >= settings.min_iterations) {
                if (residual_norm < settings.residual_tolerance_abs) {
                    result = SolverReturn::Successful;
                    break;
                }
                if (newton_iteration > 0 && residual_norm / res_0 < settings.residual_tolerance_rel) {
                    result = SolverReturn::Successful;
                    break;
                }
            }

            const int initial_cg = stats.cg_iterations;
            int64_t n_projected_now = 0;  // element Hessians projected in this Newton iteration
            bool ok = false;
            while (!ok) {
                // Project element Hessians and assemble (NewtonsMethod.cpp:254-352). Every
                // pass re-assembles from the element Hessians, projected ones included, which
                // gives the matrix that the reference's incremental update_global builds.
                bool all_projected = false;
                {
                    auto t = logger->time("project_to_PD");
                    const int32_t mirror = settings.project_to_pd_use_mirroring ? 1 : 0;
                    int64_t changed = 0;
                    switch (settings.projection_mode) {
                    case ProjectionToPD::Newton: break;
                    case ProjectionToPD::ProjectedNewton:
                    case ProjectionToPD::ProjectOnDemand:
                        if (settings.projection_mode == ProjectionToPD::ProjectedNewton || pdn_countdown > 0) {
                            check(symx_hip_problem_project(prob, settings.projection_eps, mirror, nullptr), hip, "projection");
                            all_projected = true;
                            n_projected_now = n_hessians;
                        }
                        break;
                    case ProjectionToPD::Progressive:
                        // The reference tests ppn_threshold > 0 here, so once the threshold has
                        // dropped to 0 it stops projecting and retries forever; 0 projects
                        // every block here (its "default to PN" intent).
                        if (ppn_threshold >= 0.0) {
                            if (ppn_threshold > 0.0 && ppn_threshold < 1e-12) ppn_threshold = 0.0;
                            int64_t fresh = 0, active = 0;
                            check(symx_hip_problem_project_selective(prob, settings.projection_eps, mirror, ppn_threshold, &changed, &fresh, &active),
                                  hip, "selective projection");
                            n_projected_now += fresh;
                            all_projected = active == ndofs / 3;
                        }
                        break;
                    }
                }
                {
                    auto t = logger->time("assembly");
                    // DirectLLT factors the FP64 matrix in the reference (NewtonsMethod.cpp:395-418)
                    const bool fp32_matrix = precision_mixed && settings.linear_solver == LinearSolver::BDPCG;
                    check(symx_hip_problem_assemble(prob, fp32_matrix ? SYMX_HIP_MIXED : SYMX_HIP_FP64), hip, "assembly");
                }
                ok = solve_linear_system(residual_norm);
                const bool can_project_more = settings.projection_mode != ProjectionToPD::Newton && !all_projected;
                if (!ok && !can_project_more) {
                    out->print("Linear system failed. ", Verbosity::Summary);
                    result = SolverReturn::LinearSystemSolveFailure;
                    break;
                }
                bool descends = false;
                if (ok) {
                    check(symx_hip_problem_dot(prob, SYMX_HIP_VEC_DU, SYMX_HIP_VEC_GRAD, &du_dot_grad), hip, "dot");
                    descends = du_dot_grad < 0.0;
                    if (!descends && !can_project_more) {
                        out->print("Step does not descend. ", Verbosity::Summary);
                        result = SolverReturn::StepDoesNotDescend;
                        break;
                    }
                }
                if (ok && descends) break;
                ok = false;
                out->print("Retrying with more projection.", Verbosity::Full);
                // _increase_projection (NewtonsMethod.cpp:354-368)
                if (settings.projection_mode == ProjectionToPD::ProjectOnDemand) pdn_countdown = settings.project_on_demand_countdown;
                if (settings.projection_mode == ProjectionToPD::Progressive) {
                    if (ppn_threshold < 0.0) check(symx_hip_problem_max_abs(prob, SYMX_HIP_VEC_GRAD, &ppn_threshold), hip, "max_abs");
                    ppn_threshold *= settings.ppn_tightening_factor;
                }
            }
            if (result != SolverReturn::Running) {
                out->print_with_new_line("Newton failure: Could not solve the linear system or find a descend direction.", Verbosity::Summary);
                break;
            }
            stats.cg_per_iteration.push_back(stats.cg_iterations - initial_cg);
            if (!full)
                out->print(sfmt("ph: %4.1f%% | #CG: %4d | ", n_hessians > 0 ? 100.0 * (double)n_projected_now / (double)n_hessians : 0.0,
                                stats.cg_iterations - initial_cg),
                           Verbosity::Medium);
            // _decrease_projection (NewtonsMethod.cpp:370-386)
            if (settings.projection_mode == ProjectionToPD::ProjectOnDemand) pdn_countdown--;
            if (settings.projection_mode == ProjectionToPD::Progressive) ppn_threshold *= settings.ppn_release_factor;
            stats.n_hessians += (uint64_t)n_hessians;
            stats.n_projected_hessians += (uint64_t)n_projected_now;
            logger->add_and_append("cg_iterations", last_cg_iterations);

            double du_max = 0.0;
            check(symx_hip_problem_max_abs(prob, SYMX_HIP_VEC_DU, &du_max), hip, "max_abs");
            out->print(sfmt("du: %.1e | ", du_max), Verbosity::Medium);
            if (full) out->print(sfmt("du.g: %.3e | ", du_dot_grad), Verbosity::Full);
            if (newton_iteration >= settings.min_iterations && du_max < settings.step_tolerance) {
                result = SolverReturn::Successful;
                break;
            }
            result = line_search(du_dot_grad, du_max);
            if (newton_iteration >= settings.min_iterations && !callbacks->is_converged.empty()) {
                sync_dofs_to_host();
                if (run_bool(false, callbacks->is_converged)) {
                    result = SolverReturn::Successful;
                    break;
                }
            }
            if (result != SolverReturn::Running) break;
        }

        sync_dofs_to_host();
        if (result == SolverReturn::Successful && !run_bool(true, callbacks->is_converged_state_valid)) {
            out->print_with_new_line("Newton failure: Invalid converged state.", Verbosity::Medium);
            result = SolverReturn::InvalidConvergedState;
        }
        stats.newton_iterations = newton_iteration;
        if (stats.n_hessians > 0) stats.projected_hessians_ratio = (double)stats.n_projected_hessians / (double)stats.n_hessians;
        stats.final_energy = stats.energies.empty() ? 0.0 : stats.energies.back();
        stats.final_residual = stats.residuals.empty() ? 0.0 : stats.residuals.back();
        logger->add_and_append("newton_iterations", stats.newton_iterations);
        return result;
    }

    bool NewtonsMethod::solve_linear_system(double residual_norm)
    {
        auto t = context->logger->time("linear_system_solve");
        symx_hip_ctx hip = context->hip();
        symx_hip_problem prob = dev->problem();
        check(symx_hip_problem_scale(prob, SYMX_HIP_VEC_RHS, SYMX_HIP_VEC_GRAD, -1.0), hip, "rhs");
        if (settings.linear_solver == LinearSolver::DirectLLT) {
            // DirectLLT (NewtonsMethod.cpp:395-418: Eigen SimplicialLLT of the projected FP64
            // matrix, failing when the matrix is not positive definite): the same matrix solved
            // to round-off on the device by FP64 block-Jacobi PCG, failing on p.Ap <= 0.
            symx_hip_pcg_info info{};
            const int32_t max_it = std::max<int32_t>(20 * dev->n_dofs(), 10000);
            check(symx_hip_problem_solve(prob, SYMX_HIP_FP64, 0.0, kDirectSolveTol, max_it, 1, &info), hip, "direct solve");
            stats.cg_iterations += info.n_iterations;
            last_cg_iterations = info.n_iterations;
            context->output->print(sfmt("#CG %5d | ", info.n_iterations), Verbosity::Full);
            return info.converged != 0 && info.found_indefiniteness == 0;
        }
        // Forcing sequence (NewtonsMethod.cpp:423-424).
        const double forcing = std::min(1e-2, residual_norm * std::min(0.5, std::sqrt(residual_norm)));
        const double abs_tol = std::max(forcing, settings.cg_abs_tolerance);
        symx_hip_pcg_info info{};
        check(symx_hip_problem_solve(prob, settings.precision == Precision::Mixed ? SYMX_HIP_MIXED : SYMX_HIP_FP64, abs_tol,
                                     settings.cg_rel_tolerance, settings.cg_max_iterations, settings.cg_stop_on_indefiniteness ? 1 : 0, &info),
              hip, "linear solve");
        stats.cg_iterations += info.n_iterations;
        stats.n_refinements += info.n_refinements;
        last_cg_iterations = info.n_iterations;
        return info.converged != 0;
    }

    SolverReturn NewtonsMethod::line_search(double du_dot_grad, double du_max)
    {
        symx_hip_ctx hip = context->hip();
        symx_hip_problem prob = dev->problem();
        auto& out = context->output;
        // dofs_before_ls = dofs
        check(symx_hip_problem_scale(prob, SYMX_HIP_VEC_DOFS_LS, SYMX_HIP_VEC_DOFS, 1.0), hip, "save dofs");
        auto apply = [&](double s) {
            check(symx_hip_problem_axpy(prob, SYMX_HIP_VEC_DOFS, SYMX_HIP_VEC_DOFS_LS, s, SYMX_HIP_VEC_DU), hip, "apply step");
        };
        double retraction = 1.0;
        if (du_max > settings.step_cap) {
            retraction *= settings.step_cap / du_max;
            check(symx_hip_problem_scale(prob, SYMX_HIP_VEC_DU, SYMX_HIP_VEC_DU, retraction), hip, "cap");
            out->print(sfmt("du cap %.1e | ", du_max), Verbosity::Medium);
            du_max = settings.step_cap;
            stats.ls_cap_iterations++;
        }
        double max_step = 1.0;
        if (!callbacks->max_allowed_step.empty()) {
            sync_dofs_to_host();
            for (auto& f : callbacks->max_allowed_step) max_step = std::min(max_step, f());
        }
        for (auto& f : callbacks->max_allowed_step_device) max_step = std::min(max_step, f(prob));
        if (max_step < 1.0) {
            retraction *= max_step;
            check(symx_hip_problem_scale(prob, SYMX_HIP_VEC_DU, SYMX_HIP_VEC_DU, max_step), hip, "max step");
            du_max *= max_step;
            out->print(sfmt("ls max %.1e | ", du_max), Verbosity::Medium);
            stats.ls_max_iterations++;
        }
        const double shrink = 0.5;
        double step = 1.0;
        apply(step);
        int ls_inv_it = 0;
        for (; ls_inv_it < settings.max_backtracking_invalid_state_iterations; ++ls_inv_it) {
            bool valid = true;
            if (!callbacks->is_intermediate_state_valid.empty()) {
                sync_dofs_to_host();
                for (auto& f : callbacks->is_intermediate_state_valid) valid = valid && f();
            }
            if (valid) break;
            out->print_with_new_line(sfmt("%d. step: %.2e | Invalid state", ls_inv_it, step), Verbosity::Full);
            step *= shrink;
            apply(step);
            stats.ls_inv_iterations++;
        }
        if (out->get_console_verbosity() != Verbosity::Full && ls_inv_it > 0) out->print(sfmt("ls inv %2d | ", ls_inv_it), Verbosity::Medium);
        if (ls_inv_it == settings.max_backtracking_invalid_state_iterations) {
            out->print_with_new_line("Newton failure: Too many invalid intermediate state iterations.", Verbosity::Medium);
            for (auto& f : callbacks->on_intermediate_state_invalid) f();
            return SolverReturn::TooManyInvalidIntermediateIterations;
        }
        if (!settings.enable_armijo_backtracking) {
            stats.step_sizes.push_back(step * retraction);
            return SolverReturn::Running;
        }
        // Armijo test E1 < E0 + beta * s * du.g (NewtonsMethod.cpp:565-591), evaluated as
        // E1 - E0 < beta * s * du.g with E1 - E0 summed element by element against the
        // energies of the evaluation at E0 (the last EVAL_PGH): algebraically the same test,
        // with the round-off of the element energies instead of that of the total energy.
        const double expected_decrease = settings.line_search_armijo_beta * du_dot_grad * retraction;
        double x_max = -1.0;  // max |x0|, read once when a step is short enough to need it
        auto x_scale = [&]() {
            if (x_max < 0.0) check(symx_hip_problem_max_abs(prob, SYMX_HIP_VEC_DOFS_LS, &x_max), hip, "max_abs");
            return x_max;
        };
        int armijo = 0;
        for (; armijo < settings.max_backtracking_armijo_iterations; ++armijo) {
            if (!callbacks->before_energy_evaluation.empty()) {
                sync_dofs_to_host();
                for (auto& f : callbacks->before_energy_evaluation) f();
                dev->refresh_data();
            }
            double dE = 0.0;
            {
                auto t = context->logger->time("evaluate_P");
                check(symx_hip_problem_energy_change(prob, &dE), hip, "energy");
            }
            out->print_with_new_line(sfmt("%d. step: %.2e | dE: %.3e | dE_bt: %.3e | ", armijo, step, dE, expected_decrease * step),
                                     Verbosity::Full);
            if (dE < expected_decrease * step) break;
            // Steps shorter than kTrapezoidStep * max|x|: the energy change near convergence
            // falls below the round-off of the element energies themselves (about 1e-18 of
            // each Neo-Hookean element, 5e-15 summed over 6M elements), where the test above
            // fails for every step length. There the change is estimated from directional
            // derivatives, dE ~ s (g(x0) + g(x0 + s du)) . du / 2 (trapezoid rule, error of
            // order (s |du| / element size) relative), which keeps the rounding of the
            // gradients (1e-22 here) instead.
            if (step * du_max <= kTrapezoidStep * x_scale()) {
                double g1_du = 0.0;
                check(symx_hip_problem_gradient_dot(prob, SYMX_HIP_VEC_DU, &g1_du), hip, "gradient");
                const double dE_trap = 0.5 * step * (du_dot_grad * retraction + g1_du);
                out->print(sfmt("dE (trapezoid): %.3e | ", dE_trap), Verbosity::Full);
                if (dE_trap < expected_decrease * step) {
                    stats.ls_trapezoid_accepts++;
                    break;
                }
            }
            out->print("Backtrack", Verbosity::Full);
            step *= shrink;
            apply(step);
            stats.ls_bt_iterations++;
        }
        context->logger->add_and_append("ls_bt", armijo);
        if (out->get_console_verbosity() != Verbosity::Full && armijo > 0) out->print(sfmt("ls bt %2d | ", armijo), Verbosity::Medium);
        if (armijo == settings.max_backtracking_armijo_iterations) {
            out->print_with_new_line("Newton failure: Too many armijo iterations.", Verbosity::Medium);
            for (auto& f : callbacks->on_armijo_fail) f();
            return SolverReturn::TooManyArmijoIterations;
        }
        stats.step_sizes.push_back(step * retraction);
        return SolverReturn::Running;
    }

    void NewtonsMethod::print_summary(double total_time) const
    {
        const auto& L = *context->logger;
        std::printf("\n  %-24s %10s\n", "Solve", "Total");
        std::printf("  Newton iterations        %10d\n", stats.newton_iterations);
        std::printf("  CG iterations            %10d\n", stats.cg_iterations);
        std::printf("  Line search bt           %10d\n", stats.ls_bt_iterations);
        if (total_time <= 0.0) {
            total_time = 0.0;
            for (auto& l : L.get_timer_labels()) total_time += L.get_timer_total(l);
        }
        std::printf("\n  %-40s %10s  %6s\n", "Runtime", "Time (s)", "%");
        for (auto& l : L.get_timer_labels()) {
            const double t = L.get_timer_total(l);
            std::printf("  %-40s %10.6f  %5.1f%%\n", l.c_str(), t, total_time > 0 ? 100.0 * t / total_time : 0.0);
        }
    }

    // ---- Compiled<double> (Compiled.cpp, Compilation.cpp) ----------------------------------
    namespace
    {
        bool g_check_mode = false;

        symx_hip_ctx default_ctx()
        {
            static symx_hip_ctx c = nullptr;
            if (!c) check(symx_hip_ctx_create(0, &c), nullptr, "creating the HIP context");
            return c;
        }
        std::string sha256_hex(const std::string& s)
        {
            symx_hip::Sha256 h;
            h.update(s);
            return h.hex();
        }
        // The host compiler of single-call code: SYMX_AMD_HOST_CXX, else CXX, else g++ (the
        // reference's JIT runs the system compiler the same way, Compilation.cpp:84-127).
        std::string host_compiler()
        {
            if (const char* c = std::getenv("SYMX_AMD_HOST_CXX")) return c;
            if (const char* c = std::getenv("CXX")) return c;
            return "g++";
        }
        const char* kHostFlags = "-O3 -march=native -std=c++17 -fPIC -shared";
        bool write_text(const std::string& path, const std::string& text)
        {
            const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
            std::ofstream f(tmp);
            f << text;
            f.close();
            if (!f) return false;
            return std::rename(tmp.c_str(), path.c_str()) == 0;
        }
        bool read_text(const std::string& path, std::string& text)
        {
            std::ifstream f(path);
            if (!f) return false;
            std::ostringstream ss;
            ss << f.rdbuf();
            text = ss.str();
            return true;
        }
    }

    void check_mode_ON(bool check) { g_check_mode = check; }
    bool get_check_mode_ON() { return g_check_mode; }

    void CompilationInfo::print() const
    {
        std::printf("n_inputs: %d\nn_outputs: %d\nwas_cached: %d\nruntime_codegen: %.6f s\nruntime_compilation: %.6f s\nn_bytes_symbols: %d\n", n_inputs,
                    n_outputs, was_cached ? 1 : 0, runtime_codegen, runtime_compilation, n_bytes_symbols);
    }

    Compiled<double>::Compiled(const std::vector<Scalar>& expr, const std::string& name_, const std::string& folder, const std::string& cache_id)
    {
        try_load_otherwise_compile(expr, name_, folder, cache_id);
    }
    Compiled<double>::~Compiled() { release(); }

    void Compiled<double>::release()
    {
        if (mod) symx_hip_module_unload(mod);
        mod = nullptr;
        if (lib) dlclose(lib);
        lib = nullptr;
        host_f = nullptr;
    }

    void Compiled<double>::resize_buffers()
    {
        thread_input.assign(n_threads, std::vector<double>(n_in, 0.0));
        thread_output.assign(n_threads, std::vector<double>(n_out, 0.0));
    }

    void Compiled<double>::set_n_threads(int n)
    {
        if (n < 1) throw std::runtime_error("symx: Compiled::set_n_threads needs at least one thread");
        n_threads = n;
        thread_input.resize(n, std::vector<double>(n_in, 0.0));
        thread_output.resize(n, std::vector<double>(n_out, 0.0));
    }

    // Files of one compiled version: <folder>/<name>_<sha:16>.{cpp,so,hip}; the shared object
    // exports get_sha256, which must equal the full SHA-256 of (cache_id + float type).
    bool Compiled<double>::load_files(const std::string& name_, const std::string& folder, const std::string& sha)
    {
        const std::string b = folder + "/" + name_ + "_" + sha.substr(0, 16);
        std::ifstream probe(b + ".so");
        if (!probe.good()) return false;
        void* h = dlopen((b + ".so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) return false;
        auto get_sha = reinterpret_cast<void (*)(char*)>(dlsym(h, "get_sha256"));
        auto get_ni = reinterpret_cast<int (*)()>(dlsym(h, "get_n_inputs"));
        auto get_no = reinterpret_cast<int (*)()>(dlsym(h, "get_n_outputs"));
        auto f = reinterpret_cast<void (*)(const double*, double*)>(dlsym(h, identifier(name_).c_str()));
        std::string got(64, '\0');
        if (get_sha) get_sha(&got[0]);
        std::string dsrc;
        if (!get_sha || !get_ni || !get_no || !f || got != sha || !read_text(b + ".hip", dsrc)) {
            dlclose(h);
            return false;
        }
        release();
        lib = h;
        host_f = f;
        n_in = get_ni();
        n_out = get_no();
        name = name_;
        base = b;
        device_source = dsrc;
        return true;
    }

    bool Compiled<double>::load_if_cached(const std::string& name_, const std::string& folder, const std::string& cache_id)
    {
        if (cache_id.empty()) throw std::runtime_error("symx: Compiled::load_if_cached needs a cache_id");
        const std::string dir = folder.empty() ? get_codegen_dir() : folder;
        if (!load_files(name_, dir, sha256_hex(cache_id + "0"))) return false;
        info = CompilationInfo{};
        info.n_inputs = n_in;
        info.n_outputs = n_out;
        info.was_cached = true;
        info.n_bytes_symbols = 8 * n_in;
        resize_buffers();
        return true;
    }

    void Compiled<double>::try_load_otherwise_compile(const std::vector<Scalar>& expr, const std::string& name_, const std::string& folder,
                                                      const std::string& cache_id)
    {
        if (expr.empty()) throw std::runtime_error("symx: Compiled<double> without expressions");
        // without a cache_id the content of the generated code identifies the version
        // (the reference hashes the expressions, Compilation.cpp:227-232)
        std::string id = cache_id;
        if (id.empty()) {
            Graph* g = expr[0].g;
            id = sha256_hex(emit_host_function(identifier(name_), *g, expr, g->n_symbols(), std::string(64, '0')));
        }
        if (!load_if_cached(name_, folder, id)) compile(expr, name_, folder, id);
    }

    void Compiled<double>::compile(const std::vector<Scalar>& expr, const std::string& name_, const std::string& folder, const std::string& cache_id)
    {
        if (expr.empty()) throw std::runtime_error("symx: Compiled<double> without expressions");
        const auto t0 = std::chrono::steady_clock::now();
        Graph* g = expr[0].g;
        const std::string fname = identifier(name_);
        const std::string dir = folder.empty() ? get_codegen_dir() : folder;
        const std::string id = cache_id.empty() ? sha256_hex(emit_host_function(fname, *g, expr, g->n_symbols(), std::string(64, '0'))) : cache_id;
        const std::string sha = sha256_hex(id + "0");
        const std::string host_src = emit_host_function(fname, *g, expr, g->n_symbols(), sha);
        const std::string dev_src = emit_hip_batch_function(fname, *g, expr, g->n_symbols());
        const auto t1 = std::chrono::steady_clock::now();
        mkdir(dir.c_str(), 0755);
        const std::string b = dir + "/" + name_ + "_" + sha.substr(0, 16);
        if (!write_text(b + ".cpp", host_src) || !write_text(b + ".hip", dev_src))
            throw std::runtime_error("symx: cannot write generated code to " + dir);
        const std::string so_tmp = b + ".so.tmp" + std::to_string((long long)getpid());
        const std::string cmd = host_compiler() + " " + kHostFlags + " -o \"" + so_tmp + "\" \"" + b + ".cpp\" 2>&1";
        std::string log;
        int rc = -1;
        if (FILE* pipe = popen(cmd.c_str(), "r")) {
            char buf[512];
            while (std::fgets(buf, sizeof(buf), pipe)) log += buf;
            rc = pclose(pipe);
        }
        if (rc != 0 || std::rename(so_tmp.c_str(), (b + ".so").c_str()) != 0) {
            std::remove(so_tmp.c_str());
            throw std::runtime_error("symx: host compilation of '" + name_ + "' failed (" + cmd + "):\n" + log);
        }
        const auto t2 = std::chrono::steady_clock::now();
        if (!load_files(name_, dir, sha)) throw std::runtime_error("symx: cannot load the compiled code of '" + name_ + "' from " + b + ".so");
        info = CompilationInfo{};
        info.n_inputs = n_in;
        info.n_outputs = n_out;
        info.was_cached = false;
        info.runtime_codegen = std::chrono::duration<double>(t1 - t0).count();
        info.runtime_compilation = std::chrono::duration<double>(t2 - t1).count();
        info.n_bytes_symbols = 8 * n_in;
        resize_buffers();
    }

    void Compiled<double>::set(const Scalar& s, const double& v, int thread_id)
    {
        if (!s.is_symbol()) throw std::runtime_error("symx: Compiled::set on a non-symbol");
        if (thread_id < 0 || thread_id >= n_threads) throw std::runtime_error("symx: Compiled::set thread id out of range");
        const int32_t i = s.get_symbol_idx();
        if (i >= n_in) throw std::runtime_error("symx: Compiled::set symbol outside the compiled workspace");
        thread_input[thread_id][i] = v;
    }

    void Compiled<double>::set(const Matrix& matrix, const double* v, int thread_id)
    {
        for (int32_t i = 0; i < matrix.rows(); i++)
            for (int32_t j = 0; j < matrix.cols(); j++) set(matrix(i, j), v[i * matrix.cols() + j], thread_id);
    }

    View<double> Compiled<double>::run(int thread_id)
    {
        if (!is_valid()) throw std::runtime_error("Cannot run invalid compilation for '" + name + "'");
        if (thread_id < 0 || thread_id >= n_threads) throw std::runtime_error("symx: Compiled::run thread id out of range");
        std::vector<double>& in = thread_input[thread_id];
        std::vector<double>& out = thread_output[thread_id];
        if (g_check_mode)
            for (size_t i = 0; i < in.size(); i++)
                if (std::isnan(in[i])) throw std::runtime_error("NaN detected in input #" + std::to_string(i) + " of (compiled function '" + name + "')");
        host_f(in.data(), out.data());
        if (g_check_mode)
            for (size_t i = 0; i < out.size(); i++)
                if (std::isnan(out[i])) throw std::runtime_error("NaN detected in output #" + std::to_string(i) + " of (compiled function '" + name + "')");
        return View<double>(out.data(), (int32_t)out.size());
    }

    void Compiled<double>::run_batch(const double* inputs, int64_t n, double* outputs)
    {
        if (!is_valid()) throw std::runtime_error("Cannot run invalid compilation for '" + name + "'");
        if (!mod) {
            ctx = default_ctx();
            const std::string dir = base.substr(0, base.find_last_of('/'));
            check(symx_hip_module_load(ctx, device_source.c_str(), dir.c_str(), &mod), ctx, "module load");
        }
        check(symx_hip_batch_eval(ctx, mod, identifier(name).c_str(), inputs, n, n_in, outputs, n_out), ctx, "run_batch");
    }

    // ---- host BSR ----------------------------------------------------------------------
    void BlockedSparseMatrix::to_triplets(std::vector<Triplet>& t) const
    {
        t.clear();
        for (int32_t r = 0; r < n_block_rows; r++)
            for (int64_t k = rows[r]; k < rows[r + 1]; k++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) t.push_back(Triplet{ 3 * r + i, 3 * cols[k] + j, vals[9 * k + 3 * i + j] });
    }
    std::vector<double> BlockedSparseMatrix::to_dense() const
    {
        const int64_t n = 3 * (int64_t)n_block_rows;
        std::vector<double> d((size_t)(n * n), 0.0);
        std::vector<Triplet> t;
        to_triplets(t);
        for (const Triplet& x : t) d[(size_t)x.r * n + x.c] += x.v;
        return d;
    }
    void BlockedSparseMatrix::spmxv(const double* x, double* y) const
    {
        for (int32_t r = 0; r < n_block_rows; r++) {
            double acc[3] = { 0.0, 0.0, 0.0 };
            for (int64_t k = rows[r]; k < rows[r + 1]; k++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) acc[i] += vals[9 * k + 3 * i + j] * x[3 * (int64_t)cols[k] + j];
            for (int i = 0; i < 3; i++) y[3 * (int64_t)r + i] = acc[i];
        }
    }

    // ---- ElementHessians / SecondOrderCompiledGlobal ----------------------------------
    int64_t ElementHessians::project_to_PD_inplace__all(double eps, bool mirror)
    {
        int64_t n = 0;
        check(symx_hip_problem_project(prob, eps, mirror ? 1 : 0, &n), ctx, "project_to_PD");
        return n;
    }
    spBSM ElementHessians::assemble_global(int n_threads, int ndofs)
    {
        (void)n_threads;
        check(symx_hip_problem_assemble(prob, SYMX_HIP_FP64), ctx, "assembly");
        auto m = std::make_shared<BlockedSparseMatrix>();
        m->n_block_rows = ndofs / 3;
        int64_t nnzb = 0;
        check(symx_hip_problem_export_bsr(prob, &nnzb, nullptr, nullptr, nullptr), ctx, "export_bsr");
        m->rows.resize(m->n_block_rows + 1);
        m->cols.resize(nnzb);
        m->vals.resize(9 * nnzb);
        check(symx_hip_problem_export_bsr(prob, &nnzb, m->rows.data(), m->cols.data(), m->vals.data()), ctx, "export_bsr");
        return m;
    }

    SecondOrderCompiledGlobal::SecondOrderCompiledGlobal(spGlobalPotential G_, spContext context_) : G(G_), context(context_)
    {
        dev = std::make_unique<DeviceGlobal>(G, context);
    }
    void SecondOrderCompiledGlobal::upload()
    {
        if (!dev->matches_topology()) dev = std::make_unique<DeviceGlobal>(G, context);
        dev->refresh_data();
        std::vector<double> u(dev->n_dofs());
        G->get_dofs(u.data());
        check(symx_hip_problem_set_dofs(dev->problem(), u.data()), context->hip(), "set_dofs");
    }
    void SecondOrderCompiledGlobal::evaluate_P(double& E)
    {
        upload();
        check(symx_hip_problem_eval(dev->problem(), SYMX_HIP_EVAL_P, &E), context->hip(), "evaluate_P");
    }
    void SecondOrderCompiledGlobal::evaluate_P__dP_du(double& E, std::vector<double>& grad)
    {
        upload();
        check(symx_hip_problem_eval(dev->problem(), SYMX_HIP_EVAL_PG, &E), context->hip(), "evaluate_P__dP_du");
        grad.resize(dev->n_dofs());
        check(symx_hip_problem_download(dev->problem(), SYMX_HIP_VEC_GRAD, grad.data()), context->hip(), "gradient");
    }
    spElementHessians SecondOrderCompiledGlobal::evaluate_P__dP_du__local_d2P_du2(double& E, std::vector<double>& grad)
    {
        upload();
        check(symx_hip_problem_eval(dev->problem(), SYMX_HIP_EVAL_PGH, &E), context->hip(), "evaluate_P__dP_du__local_d2P_du2");
        grad.resize(dev->n_dofs());
        check(symx_hip_problem_download(dev->problem(), SYMX_HIP_VEC_GRAD, grad.data()), context->hip(), "gradient");
        return std::make_shared<ElementHessians>(dev->problem(), context->hip());
    }
    double SecondOrderCompiledGlobal::test_derivatives_with_FD(double h)
    {
        const int n = G->get_total_n_dofs();
        std::vector<double> u(n), grad, gp, gm;
        G->get_dofs(u.data());
        double E = 0.0, Ep = 0.0, Em = 0.0;
        const spBSM H = evaluate_P__dP_du__local_d2P_du2(E, grad)->assemble_global(1, n);
        const std::vector<double> Hd = H->to_dense();
        double max_g = 0.0, max_h = 0.0;
        for (int i = 0; i < n; i++) {
            const double ui = u[i];
            u[i] = ui + h;
            G->set_dofs(u.data());
            evaluate_P__dP_du(Ep, gp);
            u[i] = ui - h;
            G->set_dofs(u.data());
            evaluate_P__dP_du(Em, gm);
            u[i] = ui;
            max_g = std::max(max_g, std::abs((Ep - Em) / (2.0 * h) - grad[i]));
            for (int j = 0; j < n; j++) max_h = std::max(max_h, std::abs((gp[j] - gm[j]) / (2.0 * h) - Hd[(size_t)i * n + j]));
        }
        G->set_dofs(u.data());
        return std::max(max_g, max_h);
    }

    // ---- CompiledInLoop<double> ---------------------------------------------------------
    CompiledInLoop<double>::CompiledInLoop(const spMWS<double>& mws_, const std::vector<Scalar>& expr, const std::string& name,
                                           const std::string& folder, const std::string& cache_id)
        : mws(mws_), kernel(identifier(name) + "_loop")
    {
        (void)cache_id;
        if (expr.empty()) throw std::runtime_error("symx: CompiledInLoop without expressions");
        if ((int32_t)mws->get_maps().size() > SYMX_HIP_MAX_MAPS) throw std::runtime_error("symx: too many data maps in one loop (max 128)");
        n_out = (int32_t)expr.size();
        KernelSpec ks;
        ks.name = kernel;
        ks.kind = KernelKind::LOOP;
        ks.outputs = expr;
        src = emit_hip_module(identifier(name), *mws->workspace().graph(), mws->get_maps(), mws->get_summation(), { ks }, symx_hip_eval_prelude());
        ctx = default_ctx();
        const std::string dir = folder.empty() ? get_codegen_dir() : folder;
        char key[80];
        int wc = 0;
        size_t cs = 0;
        check(symx_hip_jit_compile(src.c_str(), "gfx950", dir.c_str(), key, &wc, &cs), nullptr, "JIT");
        cached = wc != 0;
        check(symx_hip_module_load(ctx, src.c_str(), dir.c_str(), &mod), ctx, "module load");
    }
    CompiledInLoop<double>::~CompiledInLoop()
    {
        if (mod) symx_hip_module_unload(mod);
    }
    void CompiledInLoop<double>::run(int32_t n_threads, const Callback& cb) { run(n_threads, cb, {}); }
    void CompiledInLoop<double>::run(int32_t n_threads, const Callback& cb, const std::vector<uint8_t>& active)
    {
        (void)n_threads;
        const int32_t ne = mws->conn.n_elements();
        const int32_t stride = mws->conn.stride;
        if (!active.empty() && (int32_t)active.size() != ne) throw std::runtime_error("symx: is_element_active must have one flag per element");
        const auto& maps = mws->get_maps();
        std::vector<symx_hip_map_desc> md(maps.size());
        for (size_t m = 0; m < maps.size(); m++) {
            md[m].is_dof = 0;
            md[m].dof_offset = 0;
            md[m].host = maps[m].data();
            md[m].n_values = (int64_t)maps[m].n_elements() * maps[m].stride;
        }
        out.assign((size_t)ne * n_out, 0.0);
        const int32_t* conn = mws->conn.data();
        check(symx_hip_loop_eval(ctx, mod, kernel.c_str(), conn, ne, stride, md.data(), (int32_t)md.size(), active.empty() ? nullptr : active.data(),
                                 n_out, out.data()),
              ctx, "CompiledInLoop::run");
        for (int32_t e = 0; e < ne; e++) {
            if (!active.empty() && !active[e]) continue;
            cb(View<double>(out.data() + (size_t)e * n_out, n_out), e, 0, View<int32_t>(const_cast<int32_t*>(conn) + (size_t)e * stride, stride));
        }
    }

    // ---- FEM helpers (fem_integrators.cpp) -------------------------------------------
    std::string get_name(const FEM_Element& e)
    {
        switch (e) {
        case FEM_Element::Tet4: return "Tet4";
        case FEM_Element::Tet10: return "Tet10";
        case FEM_Element::Hex8: return "Hex8";
        case FEM_Element::Hex27: return "Hex27";
        }
        return "?";
    }
    int get_stride(const FEM_Element& e)
    {
        switch (e) {
        case FEM_Element::Tet4: return 4;
        case FEM_Element::Tet10: return 10;
        case FEM_Element::Hex8: return 8;
        case FEM_Element::Hex27: return 27;
        }
        return 0;
    }
    std::vector<std::array<double, 4>> get_integration_rule(const FEM_Element& e)
    {
        // Rules as tabulated by fem_integrators.cpp:52-120 (weight, xi0, xi1, xi2). The
        // truncated decimal constants are kept so that results match the reference.
        if (e == FEM_Element::Tet4) return { { 1.0 / 6.0, 0.25, 0.25, 0.25 } };
        if (e == FEM_Element::Tet10) {
            const double a = 0.58541020, b = 0.13819660, w = 0.041666667;
            return { { w, a, b, b }, { w, b, a, b }, { w, b, b, a }, { w, b, b, b } };
        }
        if (e == FEM_Element::Hex8) {
            const double p = 1.0 / std::sqrt(3.0);
            return { { 1.0, -p, -p, -p }, { 1.0, p, -p, -p }, { 1.0, p, p, -p }, { 1.0, -p, p, -p },
                     { 1.0, -p, -p, p }, { 1.0, p, -p, p }, { 1.0, p, p, p }, { 1.0, -p, p, p } };
        }
        // Hex27: 3-point Gauss-Legendre per axis with 10-digit truncated constants.
        const double p = 0.7745966692;
        const double wc = 0.1714677641, we = 0.2743484225, wf = 0.4389574760, wm = 0.7023319616;
        return { { wc, -p, -p, -p }, { wc, p, -p, -p }, { wc, -p, p, -p }, { wc, p, p, -p },
                 { wc, -p, -p, p }, { wc, p, -p, p }, { wc, -p, p, p }, { wc, p, p, p },
                 { we, 0.0, -p, -p }, { we, -p, 0.0, -p }, { we, p, 0.0, -p }, { we, 0.0, p, -p },
                 { we, -p, -p, 0.0 }, { we, p, -p, 0.0 }, { we, -p, p, 0.0 }, { we, p, p, 0.0 },
                 { we, 0.0, -p, p }, { we, -p, 0.0, p }, { we, p, 0.0, p }, { we, 0.0, p, p },
                 { wf, 0.0, 0.0, -p }, { wf, 0.0, -p, 0.0 }, { wf, -p, 0.0, 0.0 }, { wf, p, 0.0, 0.0 },
                 { wf, 0.0, p, 0.0 }, { wf, 0.0, 0.0, p }, { wm, 0.0, 0.0, 0.0 } };
    }

    // Node positions of Hex8 / Hex27 elements in the reference element [-1, 1]^3 as 0 (-1),
    // 1 (0), 2 (+1) per axis, in SymX's node order (fem_integrators.h:84-150: corners as
    // VTK, then mid-edges of the bottom, top and middle layers, mid-faces -x +x +y -y -z +z,
    // and the centre). The Hex27 mesh generator uses the same table.
    static const int8_t k_hex27_nodes[27][3] = {
        { 0, 0, 0 }, { 2, 0, 0 }, { 2, 2, 0 }, { 0, 2, 0 }, { 0, 0, 2 }, { 2, 0, 2 }, { 2, 2, 2 }, { 0, 2, 2 },
        { 1, 0, 0 }, { 2, 1, 0 }, { 1, 2, 0 }, { 0, 1, 0 }, { 1, 0, 2 }, { 2, 1, 2 }, { 1, 2, 2 }, { 0, 1, 2 },
        { 0, 0, 1 }, { 2, 0, 1 }, { 2, 2, 1 }, { 0, 2, 1 },
        { 0, 1, 1 }, { 2, 1, 1 }, { 1, 2, 1 }, { 1, 0, 1 }, { 1, 1, 0 }, { 1, 1, 2 },
        { 1, 1, 1 } };

    static std::vector<Scalar> shape_functions(const FEM_Element& e, const Vector& xi)
    {
        const Scalar& u = xi[0];
        const Scalar& v = xi[1];
        const Scalar& w = xi[2];
        if (e == FEM_Element::Tet4) return { 1.0 - u - v - w, u, v, w };
        if (e == FEM_Element::Tet10) {
            // quadratic tetrahedron: vertices, then edge midpoints 01 12 20 03 13 23
            const Scalar n0 = 1.0 - u - v - w;
            return { n0 * (2.0 * n0 - 1.0), u * (2.0 * u - 1.0), v * (2.0 * v - 1.0), w * (2.0 * w - 1.0),
                     4.0 * n0 * u, 4.0 * u * v, 4.0 * v * n0, 4.0 * n0 * w, 4.0 * u * w, 4.0 * v * w };
        }
        if (e == FEM_Element::Hex8) {
            // trilinear: 0.5 (1 -+ xi) per axis, corners in the order of k_hex27_nodes[0..7]
            const Scalar m[3] = { 0.5 * (1.0 - u), 0.5 * (1.0 - v), 0.5 * (1.0 - w) };
            const Scalar pl[3] = { 0.5 * (1.0 + u), 0.5 * (1.0 + v), 0.5 * (1.0 + w) };
            std::vector<Scalar> N;
            for (int n = 0; n < 8; n++) {
                const Scalar& a = k_hex27_nodes[n][0] ? pl[0] : m[0];
                const Scalar& b = k_hex27_nodes[n][1] ? pl[1] : m[1];
                const Scalar& c = k_hex27_nodes[n][2] ? pl[2] : m[2];
                N.push_back(a * b * c);
            }
            return N;
        }
        if (e == FEM_Element::Hex27) {
            // tri-quadratic Lagrange from x(x - 1), x(x + 1), 1 - x^2 per axis, scaled by 1/8
            // (corners), 1/4 (edges), 1/2 (faces), 1 (centre)
            const Scalar* X[3] = { &u, &v, &w };
            auto f = [&](int axis, int pos) -> Scalar {
                const Scalar& x = *X[axis];
                if (pos == 0) return x * (x - 1.0);
                if (pos == 2) return x * (x + 1.0);
                return 1.0 - x.powN(2);
            };
            std::vector<Scalar> N;
            for (int n = 0; n < 27; n++) {
                int n_mid = 0;
                for (int d = 0; d < 3; d++) n_mid += k_hex27_nodes[n][d] == 1;
                const double scale = n_mid == 0 ? 8.0 : (n_mid == 1 ? 4.0 : (n_mid == 2 ? 2.0 : 1.0));
                Scalar prod = f(0, k_hex27_nodes[n][0]) * f(1, k_hex27_nodes[n][1]) * f(2, k_hex27_nodes[n][2]);
                N.push_back(scale == 1.0 ? prod : prod / scale);
            }
            return N;
        }
        throw std::runtime_error("symx: shape functions for " + get_name(e) + " are not implemented");
    }

    Vector fem_interpolation(const FEM_Element& e, const std::vector<Vector>& xh, const Vector& xi)
    {
        const std::vector<Scalar> N = shape_functions(e, xi);
        if (N.size() != xh.size()) throw std::runtime_error("symx: fem_interpolation node count mismatch");
        Vector x = xh[0] * N[0];
        for (size_t i = 1; i < N.size(); i++) x = x + xh[i] * N[i];
        return x;
    }
    namespace
    {
        // Set while fem_integrator evaluates its summand: fem_jacobian of node positions at
        // the integrator's xi may then be deferred (MappedWorkspace::fem_jacobians).
        struct FemScope
        {
            MappedWorkspace<double>* mws;
            FEM_Element element;
            std::array<int32_t, 3> xi;
        };
        thread_local FemScope* g_fem_scope = nullptr;
        bool g_structured_fem = true;
    }
    void set_structured_fem(bool enabled) { g_structured_fem = enabled; }
    bool get_structured_fem() { return g_structured_fem; }

    Matrix fem_jacobian(const FEM_Element& e, const std::vector<Vector>& xh, const Vector& xi)
    {
        const Vector x = fem_interpolation(e, xh, xi);
        std::vector<Scalar> vals;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) vals.push_back(diff(x[i], xi[j]));
        FemScope* sc = g_fem_scope;
        // Elements of 8 and more nodes defer J for the structured kernels.
        if (sc && g_structured_fem && sc->element == e && get_stride(e) >= 8 && (int)xh.size() == get_stride(e) && xi.size() == 3) {
            bool ok = true;
            for (int k = 0; k < 3; k++) ok = ok && xi[k].id == sc->xi[k];
            MappedWorkspace<double>::FemJacobian rec;
            rec.n_nodes = (int32_t)xh.size();
            for (const Vector& v : xh) {
                if (v.size() != 3) ok = false;
                for (int32_t k = 0; ok && k < v.size(); k++) {
                    if (!v[k].is_symbol()) ok = false;
                    else rec.node_symbols.push_back(v[k].get_symbol_idx());
                }
            }
            if (ok) {
                Graph* g = xi[0].g;
                std::vector<Scalar> J;
                for (int k = 0; k < 9; k++) {
                    rec.expanded[k] = vals[k].id;
                    rec.J[k] = g->symbol();
                    J.push_back(Scalar(g, rec.J[k]));
                }
                sc->mws->fem_jacobians.push_back(rec);
                return Matrix(J, { 3, 3 });
            }
        }
        return Matrix(vals, { 3, 3 });
    }
    Scalar fem_integrator(MappedWorkspace<double>& mws, const FEM_Element& e, std::function<Scalar(Scalar& w, Vector& xi)> summand)
    {
        const std::vector<std::array<double, 4>> rule = get_integration_rule(e);
        return mws.add_for_each(rule, [&](Vector& ip) {
            Scalar w = ip[0];
            Vector xi({ ip[1], ip[2], ip[3] });
            FemScope scope{ &mws, e, { xi[0].id, xi[1].id, xi[2].id } };
            FemScope* prev = g_fem_scope;
            g_fem_scope = &scope;
            try {
                Scalar r = summand(w, xi);
                g_fem_scope = prev;
                return r;
            }
            catch (...) {
                g_fem_scope = prev;
                throw;
            }
        });
    }

    // ---- meshes ------------------------------------------------------------------------
    FEMMesh generate_cuboid_Tet4_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& n, const std::array<double, 3>& origin)
    {
        FEMMesh m;
        m.element_type = FEM_Element::Tet4;
        m.connectivity_stride = 4;
        const double d[3] = { size[0] / n[0], size[1] / n[1], size[2] / n[2] };
        for (int i = 0; i <= n[0]; i++)
            for (int j = 0; j <= n[1]; j++)
                for (int k = 0; k <= n[2]; k++) m.vertices.push_back({ origin[0] + i * d[0], origin[1] + j * d[1], origin[2] + k * d[2] });
        auto idx = [&](int i, int j, int k) { return (int32_t)(i * (n[1] + 1) * (n[2] + 1) + j * (n[2] + 1) + k); };
        m.connectivity.reserve((size_t)n[0] * n[1] * n[2] * 24);
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++) {
                    const int32_t v0 = idx(i, j, k), v1 = idx(i + 1, j, k), v2 = idx(i + 1, j + 1, k), v3 = idx(i, j + 1, k);
                    const int32_t v4 = idx(i, j, k + 1), v5 = idx(i + 1, j, k + 1), v6 = idx(i + 1, j + 1, k + 1), v7 = idx(i, j + 1, k + 1);
                    const int32_t t[6][4] = { { v0, v1, v6, v5 }, { v0, v1, v2, v6 }, { v0, v2, v3, v6 },
                                              { v0, v5, v6, v4 }, { v0, v6, v3, v7 }, { v0, v6, v7, v4 } };
                    for (auto& tt : t) m.connectivity.insert(m.connectivity.end(), tt, tt + 4);
                }
        m.n_elements = (int32_t)(m.connectivity.size() / 4);
        return m;
    }

    FEMMesh generate_cuboid_Tet10_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& n, const std::array<double, 3>& origin)
    {
        FEMMesh m;
        m.element_type = FEM_Element::Tet10;
        m.connectivity_stride = 10;
        const int N[3] = { 2 * n[0] + 1, 2 * n[1] + 1, 2 * n[2] + 1 };
        const double d[3] = { size[0] / (2.0 * n[0]), size[1] / (2.0 * n[1]), size[2] / (2.0 * n[2]) };
        for (int i = 0; i < N[0]; i++)
            for (int j = 0; j < N[1]; j++)
                for (int k = 0; k < N[2]; k++) m.vertices.push_back({ origin[0] + i * d[0], origin[1] + j * d[1], origin[2] + k * d[2] });
        auto idx = [&](int i, int j, int k) { return (int32_t)(i * N[1] * N[2] + j * N[2] + k); };
        // the six tetrahedra of the Tet4 block per cell (corners in the order of
        // k_hex27_nodes[0..7]); each gets its vertices, then the mid-edges 01 12 02 03 13 23,
        // the node order of the Tet10 shape functions
        static const int split[6][4] = { { 0, 1, 6, 5 }, { 0, 1, 2, 6 }, { 0, 2, 3, 6 }, { 0, 5, 6, 4 }, { 0, 6, 3, 7 }, { 0, 6, 7, 4 } };
        static const int edge[6][2] = { { 0, 1 }, { 1, 2 }, { 0, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
        m.connectivity.reserve((size_t)n[0] * n[1] * n[2] * 60);
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++)
                    for (const auto& t : split) {
                        auto node = [&](int c0, int c1) {
                            const int8_t* p = k_hex27_nodes[c0];
                            const int8_t* q = k_hex27_nodes[c1];
                            return idx(2 * i + (p[0] + q[0]) / 2, 2 * j + (p[1] + q[1]) / 2, 2 * k + (p[2] + q[2]) / 2);
                        };
                        for (int a = 0; a < 4; a++) m.connectivity.push_back(node(t[a], t[a]));
                        for (const auto& e : edge) m.connectivity.push_back(node(t[e[0]], t[e[1]]));
                    }
        m.n_elements = (int32_t)(m.connectivity.size() / 10);
        return m;
    }

    FEMMesh generate_cuboid_Hex8_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& n, const std::array<double, 3>& origin)
    {
        FEMMesh m;
        m.element_type = FEM_Element::Hex8;
        m.connectivity_stride = 8;
        const double d[3] = { size[0] / n[0], size[1] / n[1], size[2] / n[2] };
        for (int i = 0; i <= n[0]; i++)
            for (int j = 0; j <= n[1]; j++)
                for (int k = 0; k <= n[2]; k++) m.vertices.push_back({ origin[0] + i * d[0], origin[1] + j * d[1], origin[2] + k * d[2] });
        auto idx = [&](int i, int j, int k) { return (int32_t)(i * (n[1] + 1) * (n[2] + 1) + j * (n[2] + 1) + k); };
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++)
                    for (int c = 0; c < 8; c++)
                        m.connectivity.push_back(idx(i + k_hex27_nodes[c][0] / 2, j + k_hex27_nodes[c][1] / 2, k + k_hex27_nodes[c][2] / 2));
        m.n_elements = (int32_t)(m.connectivity.size() / 8);
        return m;
    }

    FEMMesh generate_cuboid_Hex27_mesh(const std::array<double, 3>& size, const std::array<int32_t, 3>& n, const std::array<double, 3>& origin)
    {
        FEMMesh m;
        m.element_type = FEM_Element::Hex27;
        m.connectivity_stride = 27;
        const int N[3] = { 2 * n[0] + 1, 2 * n[1] + 1, 2 * n[2] + 1 };
        const double d[3] = { size[0] / (2.0 * n[0]), size[1] / (2.0 * n[1]), size[2] / (2.0 * n[2]) };
        for (int i = 0; i < N[0]; i++)
            for (int j = 0; j < N[1]; j++)
                for (int k = 0; k < N[2]; k++) m.vertices.push_back({ origin[0] + i * d[0], origin[1] + j * d[1], origin[2] + k * d[2] });
        auto idx = [&](int i, int j, int k) { return (int32_t)(i * N[1] * N[2] + j * N[2] + k); };
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++)
                    for (int c = 0; c < 27; c++)
                        m.connectivity.push_back(idx(2 * i + k_hex27_nodes[c][0], 2 * j + k_hex27_nodes[c][1], 2 * k + k_hex27_nodes[c][2]));
        m.n_elements = (int32_t)(m.connectivity.size() / 27);
        return m;
    }

    FEMMesh generate_cuboid_mesh(const FEM_Element& e, const std::array<double, 3>& size, const std::array<int32_t, 3>& n, const std::array<double, 3>& origin)
    {
        switch (e) {
        case FEM_Element::Tet4: return generate_cuboid_Tet4_mesh(size, n, origin);
        case FEM_Element::Tet10: return generate_cuboid_Tet10_mesh(size, n, origin);
        case FEM_Element::Hex8: return generate_cuboid_Hex8_mesh(size, n, origin);
        case FEM_Element::Hex27: return generate_cuboid_Hex27_mesh(size, n, origin);
        }
        throw std::runtime_error("symx: generate_cuboid_mesh got an unknown element type");
    }
}
