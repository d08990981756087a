#include "compile.h"

#include "../../../include/symx_hip.h"

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace symx
{
    // ---- Element --------------------------------------------------------------------
    Index Element::operator[](int i) const
    {
        if (i < 0 || i >= size) throw std::runtime_error("symx: Element index out of range");
        return Index{ i };
    }
    Index Element::operator[](const std::string& label) const
    {
        for (size_t i = 0; i < labels.size(); i++) {
            if (labels[i] == label) return Index{ (int)i };
        }
        throw std::runtime_error("symx: Element has no label '" + label + "'");
    }
    std::vector<Index> Element::slice(int begin, int end) const
    {
        std::vector<Index> r;
        for (int i = begin; i < end; i++) r.push_back((*this)[i]);
        return r;
    }
    std::vector<Index> Element::all() const { return slice(0, size); }

    // ---- MappedWorkspace --------------------------------------------------------------
    Scalar MappedWorkspace<double>::make_scalar(std::function<std::uintptr_t()> id, std::function<const double*()> data)
    {
        maps.push_back(DataMap{ id, data, []() { return 1; }, 1, -1, ws.get_n_symbols() });
        return ws.make_scalar();
    }
    Scalar MappedWorkspace<double>::make_scalar(const double& v)
    {
        return make_scalar([&v]() { return reinterpret_cast<std::uintptr_t>(&v); }, [&v]() { return &v; });
    }
    Vector MappedWorkspace<double>::make_vector(std::function<std::uintptr_t()> id, std::function<const double*()> data, int32_t stride)
    {
        maps.push_back(DataMap{ id, data, []() { return 1; }, stride, -1, ws.get_n_symbols() });
        return ws.make_vector(stride);
    }
    Scalar MappedWorkspace<double>::make_scalar(std::function<std::uintptr_t()> id, std::function<const double*()> data,
                                                std::function<int32_t()> n, const Index& idx)
    {
        if (idx.idx < 0 || idx.idx >= conn.stride) throw std::runtime_error("symx: connectivity index out of range");
        maps.push_back(DataMap{ id, data, n, 1, idx.idx, ws.get_n_symbols() });
        return ws.make_scalar();
    }
    Vector MappedWorkspace<double>::make_vector(std::function<std::uintptr_t()> id, std::function<const double*()> data,
                                                std::function<int32_t()> n, int32_t stride, const Index& idx)
    {
        if (idx.idx < 0 || idx.idx >= conn.stride) throw std::runtime_error("symx: connectivity index out of range");
        maps.push_back(DataMap{ id, data, n, stride, idx.idx, ws.get_n_symbols() });
        return ws.make_vector(stride);
    }
    Vector MappedWorkspace<double>::make_summation_vector(const double* data, int32_t stride, int32_t n_iterations)
    {
        if (summation.n_iterations > 0) throw std::runtime_error("symx: only one summation per potential is supported");
        if (stride <= 0 || n_iterations <= 0) throw std::runtime_error("symx: empty summation table");
        summation.data.assign(data, data + (size_t)stride * n_iterations);
        summation.stride = stride;
        summation.n_iterations = n_iterations;
        summation.first_symbol_idx = ws.get_n_symbols();
        return ws.make_vector(stride);
    }
    std::vector<int32_t> MappedWorkspace<double>::maps_of_container(std::uintptr_t id) const
    {
        std::vector<int32_t> r;
        for (size_t m = 0; m < maps.size(); m++) {
            if (maps[m].id() == id) r.push_back((int32_t)m);
        }
        return r;
    }

    // ---- emitter ----------------------------------------------------------------------
    namespace
    {
        std::string literal(double v)
        {
            if (v == 0.0) return "0.0";
            char buf[64];
            std::snprintf(buf, sizeof(buf), "%a", v);
            return std::string("(") + buf + ")";
        }

        // Mark nodes reachable from the roots.
        std::vector<uint8_t> reachable(const Graph& g, const std::vector<int32_t>& roots)
        {
            std::vector<uint8_t> mark(g.size(), 0);
            std::vector<int32_t> stack(roots.begin(), roots.end());
            while (!stack.empty()) {
                const int32_t id = stack.back();
                stack.pop_back();
                if (id < 0 || mark[id]) continue;
                mark[id] = 1;
                const Node& n = g.node(id);
                switch (n.op) {
                case Op::Zero: case Op::One: case Op::Const: case Op::Symbol: break;
                case Op::PowN: stack.push_back(n.a); break;
                case Op::Branch: stack.push_back(n.a); stack.push_back(n.b); stack.push_back(n.c); break;
                default:
                    stack.push_back(n.a);
                    if (n.b >= 0) stack.push_back(n.b);
                }
            }
            return mark;
        }

        std::string ref(const Graph& g, int32_t id)
        {
            const Node& n = g.node(id);
            switch (n.op) {
            case Op::Zero: return "0.0";
            case Op::One: return "1.0";
            case Op::Const: return literal(n.value);
            case Op::Symbol: return "s" + std::to_string(n.a);
            default: return "t" + std::to_string(id);
            }
        }

        std::string expr_of(const Graph& g, int32_t id)
        {
            const Node& n = g.node(id);
            const std::string a = n.a >= 0 ? ref(g, n.a) : "";
            switch (n.op) {
            case Op::Add: return a + " + " + ref(g, n.b);
            case Op::Sub: return a + " - " + ref(g, n.b);
            case Op::Mul: return a + " * " + ref(g, n.b);
            case Op::Recip: return "1.0 / " + a;
            case Op::PowN:
                if (n.b == 2) return a + " * " + a;
                if (n.b == 3) return a + " * " + a + " * " + a;
                if (n.b == -1) return "1.0 / " + a;
                return "symx_powi(" + a + ", " + std::to_string(n.b) + ")";
            case Op::PowF: return "pow(" + a + ", " + ref(g, n.b) + ")";
            case Op::Sqrt: return "sqrt(" + a + ")";
            case Op::Ln: return "symx_ln(" + a + ")";
            case Op::Log10: return "symx_log10(" + a + ")";
            case Op::Exp: return "exp(" + a + ")";
            case Op::Sin: return "sin(" + a + ")";
            case Op::Cos: return "cos(" + a + ")";
            case Op::Tan: return "tan(" + a + ")";
            case Op::ASin: return "asin(" + a + ")";
            case Op::ACos: return "acos(" + a + ")";
            case Op::ATan: return "atan(" + a + ")";
            case Op::Branch: return "(" + ref(g, n.c) + " > 0.0) ? " + a + " : " + ref(g, n.b);
            default: throw std::runtime_error("symx: emitter met a non-operation node");
            }
        }

        bool is_op(const Node& n) { return !(n.op == Op::Zero || n.op == Op::One || n.op == Op::Const || n.op == Op::Symbol); }

        // FP32 form of a node whose operands are FP32 nodes of the same region (f<id>) or
        // FP64 values rounded at the use.
        std::string ref_f32(const Graph& g, int32_t id, const std::vector<uint8_t>& in_f32)
        {
            const Node& n = g.node(id);
            switch (n.op) {
            case Op::Zero: return "0.0f";
            case Op::One: return "1.0f";
            case Op::Const: {
                char buf[64];
                std::snprintf(buf, sizeof(buf), "%af", (double)(float)n.value);
                return std::string("(") + buf + ")";
            }
            case Op::Symbol: return "(float)s" + std::to_string(n.a);
            default: return (in_f32[id] ? "f" : "(float)t") + std::to_string(id);
            }
        }
        std::string expr_of_f32(const Graph& g, int32_t id, const std::vector<uint8_t>& in_f32)
        {
            const Node& n = g.node(id);
            const std::string a = n.a >= 0 ? ref_f32(g, n.a, in_f32) : "";
            switch (n.op) {
            case Op::Add: return a + " + " + ref_f32(g, n.b, in_f32);
            case Op::Sub: return a + " - " + ref_f32(g, n.b, in_f32);
            case Op::Mul: return a + " * " + ref_f32(g, n.b, in_f32);
            case Op::Recip: return "1.0f / " + a;
            case Op::PowN:
                if (n.b == 2) return a + " * " + a;
                if (n.b == 3) return a + " * " + a + " * " + a;
                if (n.b == -1) return "1.0f / " + a;
                return "symx_powif(" + a + ", " + std::to_string(n.b) + ")";
            case Op::PowF: return "powf(" + a + ", " + ref_f32(g, n.b, in_f32) + ")";
            case Op::Sqrt: return "sqrtf(" + a + ")";
            case Op::Ln: return "symx_lnf(" + a + ")";
            case Op::Log10: return "symx_log10f(" + a + ")";
            case Op::Exp: return "expf(" + a + ")";
            case Op::Sin: return "sinf(" + a + ")";
            case Op::Cos: return "cosf(" + a + ")";
            case Op::Tan: return "tanf(" + a + ")";
            case Op::ASin: return "asinf(" + a + ")";
            case Op::ACos: return "acosf(" + a + ")";
            case Op::ATan: return "atanf(" + a + ")";
            case Op::Branch: return "(" + ref_f32(g, n.c, in_f32) + " > 0.0f) ? " + a + " : " + ref_f32(g, n.b, in_f32);
            default: throw std::runtime_error("symx: emitter met a non-operation node");
            }
        }

        // Statements for all reachable operation nodes, in topological (id) order.
        void emit_body(std::ostringstream& os, const Graph& g, const std::vector<uint8_t>& mark, const char* indent)
        {
            for (int32_t id = 0; id < g.size(); id++) {
                if (!mark[id]) continue;
                const Node& n = g.node(id);
                if (!is_op(n)) continue;
                os << indent << "const double t" << id << " = " << expr_of(g, id) << ";\n";
            }
        }

        struct SymbolUse
        {
            std::vector<uint8_t> used;   // per symbol
            bool uses_summation = false;
        };

        SymbolUse symbol_use(const Graph& g, const std::vector<uint8_t>& mark, const SummationMap& sum)
        {
            SymbolUse u;
            u.used.assign(g.n_symbols(), 0);
            for (int32_t id = 0; id < g.size(); id++) {
                if (mark[id] && g.node(id).op == Op::Symbol) {
                    const int32_t s = g.node(id).a;
                    u.used[s] = 1;
                    if (sum.n_iterations > 0 && s >= sum.first_symbol_idx && s < sum.first_symbol_idx + sum.stride) u.uses_summation = true;
                }
            }
            return u;
        }

        void emit_gather(std::ostringstream& os, const Graph& g, const std::vector<DataMap>& maps, const SymbolUse& use,
                         const SummationMap& sum, const std::vector<std::array<int32_t, 3>>& adjoint_symbols,
                         const std::vector<int32_t>& preset_symbols = {}, const char* elem = "e")
        {
            std::vector<uint8_t> covered(g.n_symbols(), 0);
            for (int32_t s : preset_symbols) covered[s] = 1;
            std::vector<uint8_t> col_loaded(64, 0);
            for (size_t m = 0; m < maps.size(); m++) {
                const DataMap& dm = maps[m];
                bool any = false;
                for (int k = 0; k < dm.stride; k++) {
                    covered[dm.first_symbol_idx + k] = 1;
                    any = any || use.used[dm.first_symbol_idx + k];
                }
                for (const auto& a : adjoint_symbols) {
                    if (a[1] != (int32_t)m) continue;
                    covered[a[0]] = 1;
                    any = any || use.used[a[0]];
                }
                if (!any) continue;
                std::string base;
                if (dm.connectivity_index >= 0) {
                    const int c = dm.connectivity_index;
                    if (c < 64 && !col_loaded[c]) {
                        os << "        const long long i" << c << " = (long long)p.conn[" << c << " * (long long)p.n_elem + " << elem << "];\n";
                        col_loaded[c] = 1;
                    }
                    base = "i" + std::to_string(c) + " * " + std::to_string(dm.stride) + " + ";
                }
                for (int k = 0; k < dm.stride; k++) {
                    const int s = dm.first_symbol_idx + k;
                    if (!use.used[s]) continue;
                    os << "        const double s" << s << " = p.maps[" << m << "][" << base << k << "];\n";
                }
                for (const auto& a : adjoint_symbols) {
                    if (a[1] != (int32_t)m || !use.used[a[0]]) continue;
                    if (a[2] < 0 || a[2] >= dm.stride) throw std::runtime_error("symx: adjoint symbol component out of range");
                    os << "        const double s" << a[0] << " = a.lambda[" << m << "][" << base << a[2] << "];\n";
                }
            }
            if (sum.n_iterations > 0)
                for (int k = 0; k < sum.stride; k++) covered[sum.first_symbol_idx + k] = 1;
            for (int32_t s = 0; s < g.n_symbols(); s++) {
                if (use.used[s] && !covered[s]) throw std::runtime_error("symx: expression uses a symbol that is not bound to data");
            }
        }

        // Single-point gradient / gradient+Hessian kernel with chunked output staging. The
        // outputs are written in chunks of 128 bytes per row (16 doubles, or 32 floats for
        // FP32 Hessian storage): the statements are emitted segment by segment (energy and
        // the first gradient chunk, the further gradient chunks, then each Hessian chunk),
        // each segment holding the not yet emitted nodes its outputs reach, in id order; an
        // output goes to the wave's LDS staging right after its statement, and the chunk is
        // stored with coalesced writes at the end of its segment. Only values that later
        // segments reuse stay in registers, instead of every output until the end of the
        // kernel. The body runs for every lane (the element index clamped to the last
        // element, so that every gather stays in bounds); the staged rows of inactive or
        // out-of-range lanes are zeroed before each flush, and a workgroup without an active
        // element only writes zero rows.
        void emit_chunked_kernel(std::ostringstream& os, const Graph& g, const std::vector<DataMap>& maps, const SummationMap& sum,
                                 const KernelSpec& ks, bool h32)
        {
            const int n = ks.n_dofs, nb = n / 3, n_ub = nb * (nb + 1) / 2, nh = n_ub * 9, hs = hessian_record_stride(nh);
            const bool pgh = ks.kind == KernelKind::PGH;
            std::vector<int32_t> roots;
            for (const Scalar& s : ks.outputs) roots.push_back(s.id);
            const SymbolUse use = symbol_use(g, reachable(g, roots), sum);
            // output slots: gradient entries, then the Hessian's upper 3x3 blocks in storage order
            struct Chunk
            {
                bool hess;
                int c0, cw;
                std::vector<int32_t> node;  // node id of each position
            };
            std::vector<Chunk> chunks;
            for (int c0 = 0; c0 < n; c0 += 16) {
                Chunk c{ false, c0, std::min(16, n - c0), {} };
                for (int i = c0; i < c0 + c.cw; i++) c.node.push_back(ks.outputs[1 + i].id);
                chunks.push_back(c);
            }
            if (pgh) {
                std::vector<int32_t> hv;
                for (int a = 0; a < nb; a++)
                    for (int b = a; b < nb; b++)
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) hv.push_back(ks.outputs[1 + n + (3 * a + i) * n + 3 * b + j].id);
                const int cw = h32 ? 32 : 16;
                hv.resize(hs, g.zero());  // the record's pad (zeros)
                for (int c0 = 0; c0 < hs; c0 += cw) {
                    Chunk c{ true, c0, std::min(cw, hs - c0), {} };
                    c.node.assign(hv.begin() + c0, hv.begin() + c0 + c.cw);
                    chunks.push_back(c);
                }
            }
            // FP32 region of the FP32-storage twin: Hessian-only nodes between the outputs and
            // the contraction inputs (which stay FP64 and are rounded at their uses)
            std::vector<uint8_t> in_f32(g.size(), 0);
            if (pgh && h32 && !ks.f32_from.empty()) {
                std::vector<int32_t> eg_roots;
                for (int i = 0; i <= n; i++) eg_roots.push_back(ks.outputs[i].id);
                const std::vector<uint8_t> eg = reachable(g, eg_roots);
                std::vector<uint8_t> stop(g.size(), 0);
                for (int32_t id : ks.f32_from) stop[id] = 1;
                std::vector<int32_t> stack;
                for (const Chunk& c : chunks)
                    if (c.hess) stack.insert(stack.end(), c.node.begin(), c.node.end());
                std::vector<uint8_t> seen(g.size(), 0);
                while (!stack.empty()) {
                    const int32_t id = stack.back();
                    stack.pop_back();
                    if (id < 0 || seen[id]) continue;
                    seen[id] = 1;
                    const Node& nd = g.node(id);
                    if (stop[id] || eg[id] || !is_op(nd)) continue;
                    in_f32[id] = 1;
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
            os << "\nextern \"C\" __global__ void __launch_bounds__(256) " << ks.name << (h32 ? "_f32" : "") << "(const symx_hip_eval_params p)\n{\n";
            os << "    const int e = symx_block() * 256 + threadIdx.x;\n";
            os << "    const int ec = e < p.n_elem ? e : p.n_elem - 1;\n";
            os << "    const bool act = e < p.n_elem && (p.active == nullptr || p.active[ec] != 0);\n";
            os << "    const long long e0 = (long long)symx_block() * 256 + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);\n";
            os << "    const long long rem = (long long)p.n_elem - e0;\n";
            os << "    const int nv = rem < 0 ? 0 : (rem > 64 ? 64 : (int)rem);\n";
            const std::string hdst = h32 ? "(float*)p.H" : "p.H";
            os << "    if (!__syncthreads_or(act)) {\n";
            os << "        symx_zero_rows<" << n << ">(p.G, e0, nv);\n";
            if (pgh) os << "        symx_zero_rows<" << hs << ">(" << hdst << ", e0, nv);\n";
            os << "        symx_store_energy(0.0, p, e);\n        return;\n    }\n";
            os << "    double* const st = symx_stage_wave();\n";
            emit_gather(os, g, maps, use, sum, {}, {}, "ec");
            if (use.uses_summation) {  // a single summation point: its values are constants
                for (int k = 0; k < sum.stride; k++) {
                    const int s = sum.first_symbol_idx + k;
                    if (use.used[s]) os << "    const double s" << s << " = " << literal(sum.data[k]) << ";\n";
                }
            }
            std::vector<uint8_t> emitted(g.size(), 0);
            for (size_t ci = 0; ci < chunks.size(); ci++) {
                const Chunk& c = chunks[ci];
                std::vector<int32_t> seg_roots = c.node;
                if (ci == 0) seg_roots.push_back(ks.outputs[0].id);
                // positions of each node in this chunk
                std::unordered_map<int32_t, std::vector<int>> pos;
                for (int k = 0; k < c.cw; k++) pos[c.node[k]].push_back(k);
                std::vector<uint8_t> staged(c.cw, 0);
                const char* put = (c.hess && h32) ? "symx_put_f32" : "symx_put";
                // depth-first post-order from the outputs in position order: each output's
                // statements come right before it
                for (int32_t root : seg_roots) {
                    std::vector<std::pair<int32_t, int>> stack = { { root, 0 } };
                    while (!stack.empty()) {
                        const int32_t id = stack.back().first;
                        const int state = stack.back().second;
                        if (emitted[id]) {
                            stack.pop_back();
                            continue;
                        }
                        const Node& nd = g.node(id);
                        if (!is_op(nd)) {
                            emitted[id] = 1;
                            stack.pop_back();
                            continue;
                        }
                        const int32_t kids[3] = { nd.a, (nd.op == Op::PowN) ? -1 : nd.b, (nd.op == Op::Branch) ? nd.c : -1 };
                        if (state < 3) {
                            stack.back().second = state + 1;
                            if (kids[state] >= 0 && !emitted[kids[state]]) stack.push_back({ kids[state], 0 });
                            continue;
                        }
                        stack.pop_back();
                        emitted[id] = 1;
                        if (in_f32[id]) os << "    const float f" << id << " = " << expr_of_f32(g, id, in_f32) << ";\n";
                        else os << "    const double t" << id << " = " << expr_of(g, id) << ";\n";
                        auto it = pos.find(id);
                        if (it == pos.end()) continue;
                        for (int k : it->second) {
                            os << "    " << put << "(st, " << k << ", " << (in_f32[id] ? "f" : "t") << id << ");\n";
                            staged[k] = 1;
                        }
                    }
                }
                for (int k = 0; k < c.cw; k++) {
                    if (staged[k]) continue;
                    const int32_t id = c.node[k];
                    os << "    " << put << "(st, " << k << ", " << (in_f32[id] ? "f" + std::to_string(id) : ref(g, id)) << ");\n";
                }
                if (ci == 0) os << "    const double e_out = act ? " << ref(g, ks.outputs[0].id) << " : 0.0;\n";
                if (c.hess) os << "    symx_flush_rows<" << hs << ", " << c.c0 << ", " << c.cw << ">(" << hdst << ", st, e0, nv, act);\n";
                else os << "    symx_flush_rows<" << n << ", " << c.c0 << ", " << c.cw << ">(p.G, st, e0, nv, act);\n";
            }
            os << "    symx_store_energy(e_out, p, e);\n}\n";
        }
    }

    std::string emit_hip_module(const std::string& module_name, Graph& g, const std::vector<DataMap>& maps, const SummationMap& sum,
                                const std::vector<KernelSpec>& kernels, const std::string& prelude)
    {
        if (maps.size() > SYMX_HIP_MAX_MAPS) throw std::runtime_error("symx: too many data maps in one potential (max " + std::to_string(SYMX_HIP_MAX_MAPS) + ")");
        std::ostringstream os;
        os << prelude << "\n// ---- generated module: " << module_name << " ----\n";
        const bool multi_q = sum.n_iterations > 1;
        if (multi_q) {
            os << "__constant__ double " << module_name << "_sum[" << sum.data.size() << "] = {";
            for (size_t i = 0; i < sum.data.size(); i++) os << (i ? ", " : "") << literal(sum.data[i]);
            os << "};\n";
        }
        // every PGH kernel also gets a twin for FP32 Hessian storage ("_f32"): the same FP64
        // arithmetic, the Hessian rounded to float once when it is stored
        std::vector<std::pair<const KernelSpec*, bool>> todo;
        for (const KernelSpec& ks : kernels) {
            todo.push_back({ &ks, false });
            if (ks.kind == KernelKind::PGH) todo.push_back({ &ks, true });
        }
        for (const auto& item : todo) {
            const KernelSpec& ks = *item.first;
            const bool h32 = item.second;
            std::vector<int32_t> roots;
            for (const Scalar& s : ks.outputs) roots.push_back(s.id);
            if ((ks.kind == KernelKind::PG || ks.kind == KernelKind::PGH) && !multi_q) {
                emit_chunked_kernel(os, g, maps, sum, ks, h32);
                continue;
            }
            const std::vector<uint8_t> mark = reachable(g, roots);
            const SymbolUse use = symbol_use(g, mark, sum);
            const int n = ks.n_dofs;
            const int nb = n / 3;
            const int n_ub = nb * (nb + 1) / 2;

            const bool adj = ks.kind == KernelKind::PGP || ks.kind == KernelKind::LOOP;
            os << "\nextern \"C\" __global__ void __launch_bounds__(256) " << ks.name << (h32 ? "_f32" : "") << "(const symx_hip_eval_params p"
               << (adj ? ", const symx_hip_adjoint_params a" : "") << ")\n{\n";
            os << "    const int e = symx_block() * 256 + threadIdx.x;\n";
            if (ks.kind != KernelKind::C && !adj) os << "    double e_out = 0.0;\n";
            // (single-point gradient/Hessian kernels take emit_chunked_kernel above.) Summed
            // kernels accumulate their outputs in place, except the FP32-storage twins, which sum
            // the points in FP64 registers, round once and store through symx_store_rows
            const bool staged = (ks.kind == KernelKind::PG || ks.kind == KernelKind::PGH) && h32;
            if (staged) {
                os << "    double sg[" << n << "] = {};\n";
                if (ks.kind == KernelKind::PGH) os << "    double sh[" << n_ub * 9 << "] = {};\n";
            }
            os << "    if (e < p.n_elem) {\n";
            if (ks.kind != KernelKind::C) os << "        const bool act = (p.active == nullptr) || (p.active[e] != 0);\n";
            emit_gather(os, g, maps, use, sum, ks.adjoint_symbols);
            const std::string ea = "(long long)e * " + std::to_string(ks.outputs.size());
            const std::string eg = "(long long)e * " + std::to_string(n);
            const std::string eh = "(long long)e * " + std::to_string(n_ub * 9);
            // zero outputs (inactive elements, and accumulation start for multi-point sums)
            auto emit_zero = [&](const char* indent) {
                if (ks.kind == KernelKind::PG || ks.kind == KernelKind::PGH)
                    for (int i = 0; i < n; i++) os << indent << "p.G[" << eg << " + " << i << "] = 0.0;\n";
                if (ks.kind == KernelKind::PGH)
                    for (int i = 0; i < n_ub * 9; i++) os << indent << "p.H[" << eh << " + " << i << "] = 0.0;\n";
                if (adj)
                    for (size_t j = 0; j < ks.outputs.size(); j++) os << indent << "a.out[" << ea << " + " << j << "] = 0.0;\n";
            };
            if (ks.kind == KernelKind::C) {
                if (use.uses_summation) throw std::runtime_error("symx: activity conditions may not depend on summation symbols");
                emit_body(os, g, mark, "        ");
                os << "        p.C_out[e] = (" << ref(g, ks.outputs[0].id) << " > 0.0) ? 1 : 0;\n";
                os << "    }\n}\n";
                continue;
            }
            os << "        if (act) {\n";
            const char* in = "            ";
            std::string acc = "=";
            if (use.uses_summation) {
                if (multi_q) {
                    if (!staged) emit_zero(in);
                    acc = "+=";
                    os << in << "for (int q = 0; q < " << sum.n_iterations << "; ++q) {\n";
                    for (int k = 0; k < sum.stride; k++) {
                        const int s = sum.first_symbol_idx + k;
                        if (use.used[s]) os << in << "    const double s" << s << " = " << module_name << "_sum[q * " << sum.stride << " + " << k << "];\n";
                    }
                }
                else {
                    for (int k = 0; k < sum.stride; k++) {
                        const int s = sum.first_symbol_idx + k;
                        if (use.used[s]) os << in << "const double s" << s << " = " << literal(sum.data[k]) << ";\n";
                    }
                }
            }
            const std::string body_in = (use.uses_summation && multi_q) ? "                " : in;
            emit_body(os, g, mark, body_in.c_str());
            if (adj) {
                for (size_t j = 0; j < ks.outputs.size(); j++)
                    os << body_in << "a.out[" << ea << " + " << j << "] " << acc << " " << ref(g, ks.outputs[j].id) << ";\n";
                if (use.uses_summation && multi_q) os << in << "}\n";
                os << "        }\n        else {\n";
                emit_zero("            ");
                os << "        }\n    }\n}\n";
                continue;
            }
            os << body_in << "e_out += " << ref(g, ks.outputs[0].id) << ";\n";
            auto g_dst = [&](int i) { return staged ? "sg[" + std::to_string(i) + "]" : "p.G[" + eg + " + " + std::to_string(i) + "]"; };
            auto h_dst = [&](int i) { return staged ? "sh[" + std::to_string(i) + "]" : "p.H[" + eh + " + " + std::to_string(i) + "]"; };
            if (ks.kind == KernelKind::PG || ks.kind == KernelKind::PGH) {
                for (int i = 0; i < n; i++) os << body_in << g_dst(i) << " " << acc << " " << ref(g, ks.outputs[1 + i].id) << ";\n";
            }
            if (ks.kind == KernelKind::PGH) {
                int ub = 0;
                for (int a = 0; a < nb; a++) {
                    for (int b = a; b < nb; b++) {
                        for (int i = 0; i < 3; i++) {
                            for (int j = 0; j < 3; j++) {
                                const int row = 3 * a + i, col = 3 * b + j;
                                os << body_in << h_dst(ub * 9 + 3 * i + j) << " " << acc << " " << ref(g, ks.outputs[1 + n + row * n + col].id)
                                   << ";\n";
                            }
                        }
                        ub++;
                    }
                }
            }
            if (use.uses_summation && multi_q) os << in << "}\n";
            if (staged) {
                // inactive and out-of-range elements keep the zero-initialised rows
                os << "        }\n    }\n    {\n";
                os << "        const long long e0 = (long long)symx_block() * 256 + (threadIdx.x & ~63);\n";
                os << "        const long long rem = (long long)p.n_elem - e0;\n";
                os << "        const int nv = rem < 0 ? 0 : (rem > 64 ? 64 : (int)rem);\n";
                os << "        symx_store_rows<" << n << ">(p.G, sg, e0, nv);\n";
                if (ks.kind == KernelKind::PGH) os << "        symx_store_rows_f32<" << n_ub * 9 << ">((float*)p.H, sh, e0, nv);\n";
                os << "    }\n    symx_store_energy(e_out, p, e);\n}\n";
            }
            else {
                os << "        }\n        else {\n";
                emit_zero("            ");
                os << "        }\n    }\n    symx_store_energy(e_out, p, e);\n}\n";
            }
        }
        return os.str();
    }

    std::string emit_hip_fem_module(const std::string& name, Graph& g, const std::vector<DataMap>& maps, const SummationMap& sum,
                                    const FemStructure& fem, const std::string& kP, const std::string& kPG, const std::string& kPGH,
                                    const std::string& prelude)
    {
        if (maps.size() > SYMX_HIP_MAX_MAPS) throw std::runtime_error("symx: too many data maps in one potential");
        const int NN = fem.n_nodes, NQ = fem.n_q, R = (int)fem.records.size();
        const int NB = (int)fem.block_of_node.size(), NUB = (int)fem.ub_nodes.size();
        if (NB != NN || NUB != NN * (NN + 1) / 2 || (int)fem.dN.size() != NQ * NN * 3 || NQ > 64 || R < 1)
            throw std::runtime_error("symx: inconsistent FEM structure");
        std::ostringstream os;
        os << prelude << "\n// ---- generated structured FEM module: " << name << " ----\n";
        os << "__constant__ double " << name << "_dN[" << fem.dN.size() << "] = {";
        for (size_t i = 0; i < fem.dN.size(); i++) os << (i ? ", " : "") << literal(fem.dN[i]);
        os << "};\n";
        if (sum.n_iterations > 0) {
            if (sum.n_iterations != NQ) throw std::runtime_error("symx: FEM summation table does not match the quadrature");
            os << "__constant__ double " << name << "_sum[" << sum.data.size() << "] = {";
            for (size_t i = 0; i < sum.data.size(); i++) os << (i ? ", " : "") << literal(sum.data[i]);
            os << "};\n";
        }
        os << "__constant__ int " << name << "_col[" << R * NN << "] = {";
        for (int r = 0; r < R; r++)
            for (int a = 0; a < NN; a++) os << ((r || a) ? ", " : "") << fem.records[r].cols[a];
        os << "};\n__constant__ int " << name << "_blk[" << NN << "] = {";
        for (int a = 0; a < NN; a++) os << (a ? ", " : "") << fem.block_of_node[a];
        os << "};\n__constant__ int " << name << "_ub[" << 2 * NUB << "] = {";
        for (int u = 0; u < NUB; u++) os << (u ? ", " : "") << fem.ub_nodes[u][0] << ", " << fem.ub_nodes[u][1];
        os << "};\n";
        os << "__device__ __forceinline__ constexpr int " << name << "_sym(int r, int c) { return r <= c ? r * 9 - (r * (r - 1)) / 2 + (c - r) "
           << ": c * 9 - (c * (c - 1)) / 2 + (r - c); }\n";

        std::vector<int32_t> preset;
        for (const auto& rec : fem.records)
            for (int k = 0; k < 9; k++) preset.push_back(rec.J[k]);

        // kinds: 0 P, 1 PG, 2 PGH, 3 PGH with FP32 Hessian storage ("_f32")
        for (int kind4 = 0; kind4 < 4; kind4++) {
            const bool h32 = kind4 == 3;
            const int kind = h32 ? 2 : kind4;
            const std::string kname = kind == 0 ? kP : (kind == 1 ? kPG : (h32 ? kPGH + "_f32" : kPGH));
            const int NM = kind == 0 ? 1 : (kind == 1 ? 10 : 55);
            std::vector<Scalar> outs = fem.f;
            if (kind >= 1) outs.insert(outs.end(), fem.P.begin(), fem.P.end());
            if (kind == 2) outs.insert(outs.end(), fem.A.begin(), fem.A.end());
            std::vector<int32_t> roots;
            for (const Scalar& o : outs) roots.push_back(o.id);
            const std::vector<uint8_t> mark = reachable(g, roots);
            const SymbolUse use = symbol_use(g, mark, sum);

            os << "\nextern \"C\" __global__ void __launch_bounds__(256) " << kname << "(const symx_hip_eval_params p)\n{\n";
            os << "    __shared__ double s_pos[4][" << R * NN * 3 << "];\n";
            os << "    __shared__ double s_mat[4][" << NQ * NM << "];\n";
            os << "    __shared__ double s_e[4];\n";
            os << "    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;\n";
            os << "    const int e = symx_block() * 4 + wv;\n";
            os << "    const bool valid = e < p.n_elem;\n";
            os << "    const bool act = valid && (p.active == nullptr || p.active[e] != 0);\n";
            // node positions of every record into LDS
            os << "    if (act) {\n        for (int t = lane; t < " << R * NN << "; t += 64) {\n";
            os << "            const int r = t / " << NN << ";\n            const double* base = ";
            for (int r = 0; r < R - 1; r++) os << "r == " << r << " ? p.maps[" << fem.records[r].map_base << "] : ";
            os << "p.maps[" << fem.records[R - 1].map_base << "];\n";
            os << "            const long long i = (long long)p.conn[(long long)" << name << "_col[t] * p.n_elem + e];\n";
            os << "            s_pos[wv][t * 3 + 0] = base[i * 3 + 0];\n            s_pos[wv][t * 3 + 1] = base[i * 3 + 1];\n"
               << "            s_pos[wv][t * 3 + 2] = base[i * 3 + 2];\n        }\n    }\n    __syncthreads();\n";
            // material law per quadrature point (lane q)
            os << "    if (act && lane < " << NQ << ") {\n        const int q = lane;\n";
            emit_gather(os, g, maps, use, sum, {}, preset);
            if (use.uses_summation)
                for (int k = 0; k < sum.stride; k++) {
                    const int sidx = sum.first_symbol_idx + k;
                    if (use.used[sidx]) os << "        const double s" << sidx << " = " << name << "_sum[q * " << sum.stride << " + " << k << "];\n";
                }
            for (int r = 0; r < R; r++) {
                bool any = false;
                for (int k = 0; k < 9; k++) any = any || use.used[fem.records[r].J[k]];
                if (!any) continue;
                os << "        double J" << r << "[9] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };\n";
                os << "        for (int a = 0; a < " << NN << "; a++) {\n";
                os << "            const double* xa = &s_pos[wv][(" << r * NN << " + a) * 3];\n";
                os << "            const double* d = &" << name << "_dN[(q * " << NN << " + a) * 3];\n";
                os << "            #pragma unroll\n            for (int i = 0; i < 3; i++) {\n"
                   << "                J" << r << "[i * 3 + 0] += xa[i] * d[0];\n                J" << r << "[i * 3 + 1] += xa[i] * d[1];\n"
                   << "                J" << r << "[i * 3 + 2] += xa[i] * d[2];\n            }\n        }\n";
                for (int k = 0; k < 9; k++)
                    if (use.used[fem.records[r].J[k]]) os << "        const double s" << fem.records[r].J[k] << " = J" << r << "[" << k << "];\n";
            }
            emit_body(os, g, mark, "        ");
            for (int m = 0; m < NM; m++) os << "        s_mat[wv][q * " << NM << " + " << m << "] = " << ref(g, outs[m].id) << ";\n";
            os << "    }\n    __syncthreads();\n";
            // contraction
            os << "    double e_out = 0.0;\n    if (act) {\n";
            os << "        if (lane == 0)\n            for (int q = 0; q < " << NQ << "; q++) e_out += s_mat[wv][q * " << NM << "];\n";
            if (kind >= 1) {
                os << "        for (int t = lane; t < " << 3 * NN << "; t += 64) {\n";
                os << "            const int a = t / 3, i = t - 3 * a;\n            double g = 0.0;\n";
                os << "            for (int q = 0; q < " << NQ << "; q++) {\n";
                os << "                const double* d = &" << name << "_dN[(q * " << NN << " + a) * 3];\n";
                os << "                const double* P = &s_mat[wv][q * " << NM << " + 1 + i * 3];\n";
                os << "                g += P[0] * d[0] + P[1] * d[1] + P[2] * d[2];\n            }\n";
                os << "            p.G[(long long)e * " << 3 * NN << " + " << name << "_blk[a] * 3 + i] = g;\n        }\n";
            }
            if (kind == 2) {
                os << "        for (int ub = lane; ub < " << NUB << "; ub += 64) {\n";
                os << "            const int a = " << name << "_ub[2 * ub], b = " << name << "_ub[2 * ub + 1];\n";
                os << "            double h[9] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };\n";
                os << "            for (int q = 0; q < " << NQ << "; q++) {\n";
                os << "                const double* da = &" << name << "_dN[(q * " << NN << " + a) * 3];\n";
                os << "                const double* db = &" << name << "_dN[(q * " << NN << " + b) * 3];\n";
                os << "                const double* A = &s_mat[wv][q * " << NM << " + 10];\n";
                os << "                const double da0 = da[0], da1 = da[1], da2 = da[2], db0 = db[0], db1 = db[1], db2 = db[2];\n";
                os << "                #pragma unroll\n                for (int i = 0; i < 3; i++) {\n";
                os << "                    #pragma unroll\n                    for (int j = 0; j < 3; j++) {\n";
                os << "                        double acc = 0.0;\n";
                os << "                        #pragma unroll\n                        for (int k = 0; k < 3; k++) {\n";
                os << "                            const double dk = k == 0 ? da0 : (k == 1 ? da1 : da2);\n";
                os << "                            acc += dk * (A[" << name << "_sym(i * 3 + k, j * 3 + 0)] * db0 + A[" << name
                   << "_sym(i * 3 + k, j * 3 + 1)] * db1 + A[" << name << "_sym(i * 3 + k, j * 3 + 2)] * db2);\n";
                os << "                        }\n                        h[i * 3 + j] += acc;\n                    }\n                }\n            }\n";
                if (h32) {
                    os << "            float* dst = (float*)p.H + (long long)e * " << NUB * 9 << " + ub * 9;\n";
                    os << "            #pragma unroll\n            for (int k = 0; k < 9; k++) dst[k] = (float)h[k];\n        }\n";
                }
                else {
                    os << "            double* dst = p.H + (long long)e * " << NUB * 9 << " + ub * 9;\n";
                    os << "            #pragma unroll\n            for (int k = 0; k < 9; k++) dst[k] = h[k];\n        }\n";
                }
            }
            os << "    }\n";
            if (kind >= 1) {
                os << "    else if (valid) {\n";
                os << "        for (int t = lane; t < " << 3 * NN << "; t += 64) p.G[(long long)e * " << 3 * NN << " + t] = 0.0;\n";
                if (kind == 2)
                    os << "        for (int t = lane; t < " << NUB * 9 << "; t += 64) " << (h32 ? "((float*)p.H)" : "p.H") << "[(long long)e * " << NUB * 9
                       << " + t] = 0.0;\n";
                os << "    }\n";
            }
            os << "    if (lane == 0) s_e[wv] = symx_element_energy(e_out, p, e);\n    __syncthreads();\n";
            os << "    if (threadIdx.x == 0) p.E_partial[symx_block()] = ((s_e[0] + s_e[1]) + s_e[2]) + s_e[3];\n}\n";
        }
        return os.str();
    }

    std::string emit_hip_batch_function(const std::string& name, Graph& g, const std::vector<Scalar>& outputs, int32_t n_inputs)
    {
        std::vector<int32_t> roots;
        for (const Scalar& s : outputs) roots.push_back(s.id);
        const std::vector<uint8_t> mark = reachable(g, roots);
        std::ostringstream os;
        os << "__device__ __forceinline__ double symx_ln(double x) { return (x <= 0.0) ? -__builtin_inf() : log(x); }\n"
           << "__device__ __forceinline__ double symx_log10(double x) { return (x <= 0.0) ? -__builtin_inf() : log10(x); }\n"
           << "__device__ __forceinline__ double symx_powi(double x, int n) { const bool ng = n < 0; unsigned m = ng ? -n : n; double r = 1.0, b = x;"
              " while (m) { if (m & 1u) r *= b; b *= b; m >>= 1; } return ng ? 1.0 / r : r; }\n";
        os << "extern \"C\" __global__ void __launch_bounds__(256) " << name << "(const double* in, double* out, int n)\n{\n";
        os << "    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;\n    if (i >= n) return;\n";
        for (int32_t id = 0; id < g.size(); id++) {
            if (mark[id] && g.node(id).op == Op::Symbol) {
                const int s = g.node(id).a;
                if (s >= n_inputs) throw std::runtime_error("symx: symbol outside the input range");
                os << "    const double s" << s << " = in[i * " << n_inputs << " + " << s << "];\n";
            }
        }
        emit_body(os, g, mark, "    ");
        for (size_t k = 0; k < outputs.size(); k++) os << "    out[i * " << outputs.size() << " + " << k << "] = " << ref(g, outputs[k].id) << ";\n";
        os << "}\n";
        return os.str();
    }

    std::string emit_host_function(const std::string& name, Graph& g, const std::vector<Scalar>& outputs, int32_t n_inputs,
                                   const std::string& sha256_hex)
    {
        std::vector<int32_t> roots;
        for (const Scalar& s : outputs) roots.push_back(s.id);
        const std::vector<uint8_t> mark = reachable(g, roots);
        std::ostringstream os;
        os << "// generated by symx_amd: " << name << "\n#include <cmath>\n#include <cstring>\nusing namespace std;\n"
           << "static inline double symx_ln(double x) { return (x <= 0.0) ? -INFINITY : log(x); }\n"
           << "static inline double symx_log10(double x) { return (x <= 0.0) ? -INFINITY : log10(x); }\n"
           << "static inline double symx_powi(double x, int n) { const bool ng = n < 0; unsigned m = ng ? -n : n; double r = 1.0, b = x;"
              " while (m) { if (m & 1u) r *= b; b *= b; m >>= 1; } return ng ? 1.0 / r : r; }\n";
        os << "extern \"C\" void " << name << "(const double* in, double* out)\n{\n";
        for (int32_t id = 0; id < g.size(); id++) {
            if (mark[id] && g.node(id).op == Op::Symbol) {
                const int s = g.node(id).a;
                if (s >= n_inputs) throw std::runtime_error("symx: symbol outside the input range");
                os << "    const double s" << s << " = in[" << s << "];\n";
            }
        }
        emit_body(os, g, mark, "    ");
        for (size_t k = 0; k < outputs.size(); k++) os << "    out[" << k << "] = " << ref(g, outputs[k].id) << ";\n";
        os << "}\n";
        os << "extern \"C\" int get_n_inputs() { return " << n_inputs << "; }\n";
        os << "extern \"C\" int get_n_outputs() { return " << outputs.size() << "; }\n";
        os << "extern \"C\" int get_float_type() { return 0; }\n";  // FloatType::Double
        os << "extern \"C\" void get_sha256(char* out) { std::memcpy(out, \"" << sha256_hex << "\", 64); }\n";
        return os.str();
    }

    int32_t hessian_record_stride(int32_t n_values) { return n_values > 16 ? (n_values + 95) / 96 * 96 : n_values; }

    int64_t count_operations(Graph& g, const std::vector<Scalar>& outputs)
    {
        std::vector<int32_t> roots;
        for (const Scalar& s : outputs) roots.push_back(s.id);
        const std::vector<uint8_t> mark = reachable(g, roots);
        int64_t c = 0;
        for (int32_t id = 0; id < g.size(); id++)
            if (mark[id] && is_op(g.node(id))) c++;
        return c;
    }
}
