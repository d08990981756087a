// Mapped workspaces (symbol <-> user data binding) and the HIP emitter.
//
// Mirrors SymX's symx/src/compile/{MappedWorkspace.h,data_maps.h,Element.h,
// LabelledConnectivity.h}: potentials declare symbols bound to user arrays
// through a connectivity table, and the emitter turns the resulting
// expressions into element kernels for gfx950 (replacing Compilation.cpp's
// C++/AVX2 emitter, Compilation.cpp:243-615).
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "symbol.h"

namespace symx
{
    // Element stencil index (Element.h).
    struct Index { int idx; };

    struct Element
    {
        explicit Element(int n_items) : size(n_items) {}
        Index operator[](int i) const;
        Index operator[](const std::string& label) const;
        std::vector<Index> slice(int begin, int end) const;
        std::vector<Index> all() const;
        void set_labels(const std::vector<std::string>& l) { labels = l; }
        int n_items() const { return size; }

    private:
        int size = -1;
        std::vector<std::string> labels;
    };

    // Connectivity table with named columns (LabelledConnectivity.h).
    template <std::size_t N>
    struct LabelledConnectivity
    {
        std::array<std::string, N> labels;
        std::vector<std::array<int32_t, N>> data;
        LabelledConnectivity(const std::array<std::string, N>& l) : labels(l) {}
        Element get_element() const
        {
            Element e((int)N);
            e.set_labels(std::vector<std::string>(labels.begin(), labels.end()));
            return e;
        }
        void push_back(const std::array<int32_t, N>& row) { data.push_back(row); }
        // Prepend the running element index as the first column.
        void numbered_push_back(const std::array<int32_t, N - 1>& row)
        {
            std::array<int32_t, N> r;
            r[0] = (int32_t)data.size();
            for (std::size_t i = 0; i + 1 < N; i++) r[i + 1] = row[i];
            data.push_back(r);
        }
        std::size_t size() const { return data.size(); }
    };

    // Connectivity (data_maps.h:29-70): lambdas so that user arrays may resize.
    struct ConnectivityMap
    {
        std::function<const int32_t*()> data;
        std::function<int32_t()> n_elements;
        int32_t stride = -1;
        ConnectivityMap() = default;
        ConnectivityMap(std::function<const int32_t*()> d, std::function<int32_t()> n, int32_t s) : data(d), n_elements(n), stride(s) {}
        ConnectivityMap(const std::vector<int32_t>& arr, int32_t s)
            : data([&arr]() { return arr.data(); }), n_elements([&arr, s]() { return (int32_t)(arr.size() / s); }), stride(s) {}
        template <std::size_t N>
        ConnectivityMap(const std::vector<std::array<int32_t, N>>& arr)
            : data([&arr]() { return arr.empty() ? nullptr : arr[0].data(); }), n_elements([&arr]() { return (int32_t)arr.size(); }),
              stride((int32_t)N) {}
    };

    // Binding of `stride` consecutive symbols to user data (data_maps.h:86-105).
    struct DataMap
    {
        std::function<std::uintptr_t()> id;       // container identity (DoF detection)
        std::function<const double*()> data;
        std::function<int32_t()> n_elements;
        int32_t stride = -1;
        int32_t connectivity_index = -1;          // -1: same value for all elements
        int32_t first_symbol_idx = -1;
        int32_t size() const { return n_elements() * stride; }
    };

    // Symbols that iterate over a fixed table (quadrature points) and are
    // summed over (MappedWorkspace::make_summation_vector / add_for_each).
    struct SummationMap
    {
        std::vector<double> data;  // [n_iterations][stride]
        int32_t stride = 0;
        int32_t n_iterations = 0;
        int32_t first_symbol_idx = -1;
    };

    template <typename FLOAT>
    class MappedWorkspace;

    template <>
    class MappedWorkspace<double>
    {
    public:
        explicit MappedWorkspace(ConnectivityMap c) : conn(c) {}
        static std::pair<std::shared_ptr<MappedWorkspace<double>>, Element> create(ConnectivityMap c)
        {
            auto m = std::make_shared<MappedWorkspace<double>>(c);
            return { m, Element(c.stride) };
        }
        static std::pair<std::shared_ptr<MappedWorkspace<double>>, Element> create(const std::vector<int32_t>& arr, int32_t stride)
        {
            return create(ConnectivityMap(arr, stride));
        }
        template <std::size_t N>
        static std::pair<std::shared_ptr<MappedWorkspace<double>>, Element> create(const std::vector<std::array<int32_t, N>>& arr)
        {
            return create(ConnectivityMap(arr));
        }
        template <std::size_t N>
        static std::pair<std::shared_ptr<MappedWorkspace<double>>, Element> create(const LabelledConnectivity<N>& c)
        {
            auto m = std::make_shared<MappedWorkspace<double>>(ConnectivityMap(c.data));
            return { m, c.get_element() };
        }

        // Fixed (per potential) values
        Scalar make_scalar(std::function<std::uintptr_t()> id, std::function<const double*()> data);
        Scalar make_scalar(const double& v);
        Vector make_vector(std::function<std::uintptr_t()> id, std::function<const double*()> data, int32_t stride);
        template <typename STATIC_VECTOR>
        Vector make_vector(const STATIC_VECTOR& arr)
        {
            return make_vector([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return (const double*)arr.data(); },
                               (int32_t)arr.size());
        }

        // Per element values through the connectivity
        Scalar make_scalar(std::function<std::uintptr_t()> id, std::function<const double*()> data, std::function<int32_t()> n, const Index& idx);
        template <typename DYNAMIC_VECTOR>
        Scalar make_scalar(const DYNAMIC_VECTOR& arr, const Index& idx)
        {
            return make_scalar([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return (const double*)arr.data(); },
                               [&arr]() { return (int32_t)arr.size(); }, idx);
        }
        template <typename DYNAMIC_VECTOR>
        std::vector<Scalar> make_scalars(const DYNAMIC_VECTOR& arr, const std::vector<Index>& indices)
        {
            std::vector<Scalar> r;
            for (const auto& i : indices) r.push_back(make_scalar(arr, i));
            return r;
        }
        Vector make_vector(std::function<std::uintptr_t()> id, std::function<const double*()> data, std::function<int32_t()> n, int32_t stride,
                           const Index& idx);
        template <typename STATIC_VECTOR>
        Vector make_vector(const std::vector<STATIC_VECTOR>& arr, const Index& idx)
        {
            constexpr int32_t stride = (int32_t)(sizeof(STATIC_VECTOR) / sizeof(double));
            return make_vector([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); },
                               [&arr]() { return reinterpret_cast<const double*>(arr.data()); }, [&arr]() { return (int32_t)arr.size(); },
                               stride, idx);
        }
        template <typename DYNAMIC_VECTOR>
        Vector make_vector(const DYNAMIC_VECTOR& arr, int32_t stride, const Index& idx)
        {
            return make_vector([&arr]() { return reinterpret_cast<std::uintptr_t>(&arr); }, [&arr]() { return (const double*)arr.data(); },
                               [&arr, stride]() { return (int32_t)(arr.size() / stride); }, stride, idx);
        }
        template <typename STATIC_VECTOR>
        std::vector<Vector> make_vectors(const std::vector<STATIC_VECTOR>& arr, const std::vector<Index>& indices)
        {
            std::vector<Vector> r;
            for (const auto& i : indices) r.push_back(make_vector(arr, i));
            return r;
        }
        template <typename STATIC_VECTOR>
        std::vector<Vector> make_vectors(const std::vector<STATIC_VECTOR>& arr, const Element& elem)
        {
            return make_vectors(arr, elem.all());
        }
        template <typename DYNAMIC_VECTOR>
        std::vector<Vector> make_vectors(const DYNAMIC_VECTOR& arr, int32_t stride, const std::vector<Index>& indices)
        {
            std::vector<Vector> r;
            for (const auto& i : indices) r.push_back(make_vector(arr, stride, i));
            return r;
        }

        // Per element matrices (row-major, shape[0] x shape[1] doubles per item)
        template <typename STATIC_MATRIX>
        Matrix make_matrix(const std::vector<STATIC_MATRIX>& arr, const std::array<int32_t, 2>& shape, const Index& idx)
        {
            constexpr int32_t stride = (int32_t)(sizeof(STATIC_MATRIX) / sizeof(double));
            if (stride != shape[0] * shape[1]) throw std::runtime_error("symx: make_matrix shape does not match the item size");
            Vector v = make_vector(arr, idx);
            return Matrix(v.values(), shape);
        }

        // Summation over a table of iteration vectors (quadrature).
        Vector make_summation_vector(const double* data, int32_t stride, int32_t n_iterations);
        template <typename STATIC_VECTOR>
        Scalar add_for_each(const std::vector<STATIC_VECTOR>& iteration_vectors, std::function<Scalar(Vector&)> f)
        {
            constexpr int32_t stride = (int32_t)(sizeof(STATIC_VECTOR) / sizeof(double));
            Vector v = make_summation_vector(reinterpret_cast<const double*>(iteration_vectors.data()), stride, (int32_t)iteration_vectors.size());
            return f(v);
        }

        Scalar make_zero() { return ws.get_zero(); }
        Scalar make_one() { return ws.get_one(); }
        Vector make_zero_vector(int32_t n) { return ws.get_zero_vector(n); }
        Matrix make_zero_matrix(const std::array<int32_t, 2>& s) { return ws.get_zero_matrix(s); }
        Matrix make_identity_matrix(int32_t n) { return ws.get_identity_matrix(n); }
        Matrix get_identity_matrix(int32_t n) { return ws.get_identity_matrix(n); }

        // Introspection
        Workspace& workspace() { return ws; }
        const std::vector<DataMap>& get_maps() const { return maps; }
        const SummationMap& get_summation() const { return summation; }
        bool has_summation() const { return summation.n_iterations > 0; }
        ConnectivityMap conn;

        // Symbols bound to a DoF container, in map creation order (DoF identification of
        // SecondOrderCompiledPotential.cpp:10-33 via MappedWorkspace.h:800-825).
        std::vector<int32_t> maps_of_container(std::uintptr_t id) const;

        // Isoparametric Jacobians that fem_jacobian deferred inside fem_integrator: the
        // expression uses the 9 symbols J (row-major, J_ik = d x_i / d xi_k), and J equals
        // `expanded` = sum over nodes of x_a (x) dN_a/dxi. Code generation either keeps J
        // (structured FEM kernels) or substitutes `expanded` back.
        struct FemJacobian
        {
            int32_t n_nodes = 0;
            std::vector<int32_t> node_symbols;  // [n_nodes][3] symbol ids of the node positions
            std::array<int32_t, 9> J{};         // graph nodes of the deferred symbols
            std::array<int32_t, 9> expanded{};  // graph nodes of the expansion
        };
        std::vector<FemJacobian> fem_jacobians;

    private:
        Workspace ws;
        std::vector<DataMap> maps;
        SummationMap summation;
    };

    template <typename FLOAT>
    using spMWS = std::shared_ptr<MappedWorkspace<FLOAT>>;

    // ---- HIP emitter --------------------------------------------------------------
    // Description of one generated element kernel.
    // P: energy; PG: + gradient; PGH: + Hessian; C: activity condition; PGP: adjoint mixed
    // partials; LOOP: arbitrary per-element outputs (CompiledInLoop). PGP and LOOP take a
    // second symx_hip_adjoint_params argument and write its `out` array.
    enum class KernelKind { P, PG, PGH, C, PGP, LOOP };
    struct KernelSpec
    {
        std::string name;
        KernelKind kind = KernelKind::P;
        std::vector<Scalar> outputs;   // P: [E]; PG: [E, g]; PGH: [E, g, H row-major n x n]; C: [cond]; PGP: [d(lambda.g)/dp_j]
        int32_t n_dofs = 0;            // DoFs of the element (multiple of 3 for PG/PGH)
        // PGP only: adjoint symbols {symbol id, data map m, component k}. Symbol `id`
        // reads the adjoint vector where DoF map m reads its component k.
        std::vector<std::array<int32_t, 3>> adjoint_symbols;
        // PGH with derivatives through a linear cut: the FP32-storage twin forms the Hessian
        // contraction in FP32, i.e. the Hessian-only nodes between the outputs and these nodes
        // (the cut's second derivatives and the coefficients dL/dx, computed in FP64 and
        // rounded once); empty: the whole kernel computes in FP64.
        std::vector<int32_t> f32_from;
    };

    // Emit one hiprtc translation unit holding all kernels of a potential. The
    // gather follows `maps` (one map per entry of the eval params), summation
    // symbols are baked as constant tables, and outputs follow the layout of
    // symx_hip_eval_params.
    std::string emit_hip_module(const std::string& module_name, Graph& g, const std::vector<DataMap>& maps, const SummationMap& sum,
                                const std::vector<KernelSpec>& kernels, const std::string& prelude);

    // Structured FEM element kernels (one wavefront per element, four per workgroup):
    // the expression depends on the DoFs only through the deferred Jacobian of record 0,
    //   E_e = sum_q f(J_q, ...),  J_q = sum_a x_a (x) dN_a(xi_q),
    // so g_a = sum_q P_q dN_a and H_ab = sum_q dN_a^T A_q dN_b with P = df/dJ and
    // A = d2f/dJ2 generated symbolically (9 and 45 outputs) and the contraction written
    // by hand. Further records (e.g. rest positions) are data Jacobians computed the same way.
    struct FemStructure
    {
        struct Record
        {
            int32_t map_base = -1;             // data map whose array holds the node positions
            std::vector<int32_t> cols;         // connectivity column per node
            std::array<int32_t, 9> J{};        // symbol ids of the deferred Jacobian
        };
        int32_t n_nodes = 0;
        int32_t n_q = 0;
        std::vector<double> dN;                // [n_q][n_nodes][3]
        std::vector<Record> records;           // records[0]: DoF positions
        std::vector<int32_t> block_of_node;    // DoF block of each node
        std::vector<std::array<int32_t, 2>> ub_nodes;  // nodes of each upper block (A <= B, block order)
        std::vector<Scalar> f, P, A;           // f; df/dJ (9); d2f/dJ2 upper triangle (45)
    };
    std::string emit_hip_fem_module(const std::string& module_name, Graph& g, const std::vector<DataMap>& maps, const SummationMap& sum,
                                    const FemStructure& fem, const std::string& kP, const std::string& kPG, const std::string& kPGH,
                                    const std::string& prelude);

    // Emit a standalone kernel evaluating `outputs` for a batch of input vectors
    // (Compiled<double>): in [n][n_inputs], out [n][n_outputs].
    std::string emit_hip_batch_function(const std::string& name, Graph& g, const std::vector<Scalar>& outputs, int32_t n_inputs);
    // Host C++ translation unit of the same expressions for single host calls (the reference's
    // own JIT target, Compilation.cpp:243-336): extern "C" void name(const double* in, double* out),
    // get_n_inputs(), get_n_outputs(), get_float_type() and get_sha256(char* 64 hex digits).
    std::string emit_host_function(const std::string& name, Graph& g, const std::vector<Scalar>& outputs, int32_t n_inputs,
                                   const std::string& sha256_hex);

    // Values per element Hessian record written by the single-point (chunked) PGH kernels
    // for a Hessian of n_values upper-block values: records longer than one 128-byte chunk
    // of doubles are padded to a multiple of 96 values: a multiple of 32, so that every chunk
    // the kernels store (32 floats or 16 doubles) is one aligned 128-byte line, and of 3, as the
    // runtime addresses records in 3-value units (2 and 3 blocks: 27 and 54 values -> 96; 4 blocks:
    // 90 -> 96).
    int32_t hessian_record_stride(int32_t n_values);

    // Number of SSA operations that a kernel with these outputs executes (for reports).
    int64_t count_operations(Graph& g, const std::vector<Scalar>& outputs);
}
