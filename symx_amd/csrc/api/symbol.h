// Symbolic expression layer: a hash-consed DAG of scalar operations with
// symbolic differentiation. Mirrors the user-facing surface of SymX's
// symx/src/symbol/{Scalar,Vector,Matrix,Workspace,diff}.h so that potential
// definitions written against SymX compile unchanged against this library.
//
// Reference behaviour followed (SymX, path:line):
//   - node kinds and the Ln/Log10 guard:        symx/src/symbol/Expr.h:11-38
//   - constants within 10*eps snap to 0 / 1:      symx/src/symbol/Expressions.cpp:41-60
//   - CSE by (type, a, b, cond) key:             symx/src/symbol/Expressions.cpp:232-275
//   - derivative rules:                          symx/src/symbol/diff.cpp:7-84
//   - symmetric Hessian mirrors upper triangle:  symx/src/symbol/diff.cpp:145-186
//   - 3x3 det / inverse via cofactors:           symx/src/symbol/Matrix.cpp:175-430
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace symx
{
    enum class Op : int32_t
    {
        Zero = 0, One, Const, Symbol,
        Add, Sub, Mul, Recip, PowN, PowF, Sqrt, Ln, Log10, Exp,
        Sin, Cos, Tan, ASin, ACos, ATan,
        Branch,   // c > 0 ? a : b
        Count
    };
    const char* op_name(Op op);

    struct Node
    {
        Op op = Op::Zero;
        int32_t a = -1;     // first operand, symbol id for Symbol
        int32_t b = -1;     // second operand, integer exponent for PowN
        int32_t c = -1;     // condition operand for Branch
        double value = 0.0; // Const payload
    };

    // Expression graph. Node ids are topologically ordered: every operand has a
    // smaller id than the node that uses it.
    class Graph
    {
    public:
        Graph();
        int32_t zero() const { return 0; }
        int32_t one() const { return 1; }
        int32_t constant(double v);
        int32_t symbol();
        int32_t op(Op op, int32_t a, int32_t b = -1);
        int32_t branch(int32_t cond, int32_t pos, int32_t neg);

        const Node& node(int32_t id) const { return nodes[id]; }
        int32_t size() const { return (int32_t)nodes.size(); }
        int32_t n_symbols() const { return (int32_t)symbol_nodes.size(); }
        int32_t symbol_node(int32_t symbol_id) const { return symbol_nodes[symbol_id]; }

        // Reference evaluator (host, FP64) used by Compiled<double> checks and tests.
        void set_symbol_value(int32_t symbol_id, double v) { symbol_values[symbol_id] = v; }
        double get_symbol_value(int32_t symbol_id) const { return symbol_values[symbol_id]; }
        double eval(int32_t id) const;
        std::vector<double> eval_many(const std::vector<int32_t>& ids) const;

        // Canonical text of the sub-DAG reachable from `roots` (ids renumbered in
        // visit order) - used as a content key for the code-object cache.
        std::string canonical(const std::vector<int32_t>& roots) const;

        // Rebuild the sub-DAGs of `roots` with the nodes in `replace` (node id -> node id)
        // substituted, through op()/branch() so that folding and CSE apply; returns the
        // new roots.
        std::vector<int32_t> substitute(const std::vector<int32_t>& roots, const std::unordered_map<int32_t, int32_t>& replace);

    private:
        struct KeyHash { size_t operator()(const std::array<int64_t, 4>& k) const; };
        int32_t find_or_add(const Node& n);
        std::vector<Node> nodes;
        std::vector<int32_t> symbol_nodes;
        std::vector<double> symbol_values;
        std::unordered_map<std::array<int64_t, 4>, int32_t, KeyHash> cse;
    };

    class Scalar
    {
    public:
        Graph* g = nullptr;
        int32_t id = -1;

        Scalar() = default;
        Scalar(Graph* g, int32_t id) : g(g), id(id) {}

        Scalar operator+(const Scalar& o) const;
        Scalar operator-(const Scalar& o) const;
        Scalar operator*(const Scalar& o) const;
        Scalar operator/(const Scalar& o) const;
        Scalar operator+(double v) const;
        Scalar operator-(double v) const;
        Scalar operator*(double v) const;
        Scalar operator/(double v) const;
        void operator+=(const Scalar& o) { *this = *this + o; }
        void operator-=(const Scalar& o) { *this = *this - o; }
        void operator*=(const Scalar& o) { *this = *this * o; }
        void operator/=(const Scalar& o) { *this = *this / o; }
        void operator+=(double v) { *this = *this + v; }
        void operator-=(double v) { *this = *this - v; }
        void operator*=(double v) { *this = *this * v; }
        void operator/=(double v) { *this = *this / v; }

        Scalar powN(int32_t n) const;
        Scalar powF(double p) const;
        Scalar inv() const;
        Scalar sqrt() const;
        Scalar log() const;
        Scalar ln() const;
        Scalar log10() const;
        Scalar exp() const;
        Scalar sin() const;
        Scalar cos() const;
        Scalar tan() const;
        Scalar asin() const;
        Scalar acos() const;
        Scalar atan() const;

        bool is_zero() const { return id == 0; }
        bool is_one() const { return id == 1; }
        bool is_symbol() const { return g->node(id).op == Op::Symbol; }
        bool is_constant() const;
        double constant_value() const;
        int32_t get_symbol_idx() const { return g->node(id).a; }
        void set_value(double v) { g->set_symbol_value(get_symbol_idx(), v); }
        double get_value() const { return g->get_symbol_value(get_symbol_idx()); }
        double eval() const { return g->eval(id); }

        Scalar get_zero() const { return Scalar(g, 0); }
        Scalar get_one() const { return Scalar(g, 1); }
        Scalar make_constant(double v) const { return Scalar(g, g->constant(v)); }
    };

    Scalar operator+(double v, const Scalar& s);
    Scalar operator-(double v, const Scalar& s);
    Scalar operator*(double v, const Scalar& s);
    Scalar operator/(double v, const Scalar& s);
    Scalar operator-(const Scalar& s);
    Scalar powN(const Scalar& s, int32_t n);
    Scalar powF(const Scalar& s, double p);
    Scalar sqrt(const Scalar& s);
    Scalar log(const Scalar& s);
    Scalar ln(const Scalar& s);
    Scalar log10(const Scalar& s);
    Scalar exp(const Scalar& s);
    Scalar sin(const Scalar& s);
    Scalar cos(const Scalar& s);
    Scalar tan(const Scalar& s);
    Scalar asin(const Scalar& s);
    Scalar acos(const Scalar& s);
    Scalar atan(const Scalar& s);

    // Branch: evaluates to `pos` where cond > 0 and to `neg` otherwise.
    Scalar branch(const Scalar& cond, const Scalar& pos, const Scalar& neg);
    Scalar branch(const Scalar& cond, double pos, const Scalar& neg);
    Scalar branch(const Scalar& cond, const Scalar& pos, double neg);
    Scalar branch(const Scalar& cond, double pos, double neg);
    Scalar operator>(const Scalar& a, const Scalar& b);  // 1 where a > b, else 0
    Scalar operator<(const Scalar& a, const Scalar& b);
    Scalar operator>(const Scalar& a, double b);
    Scalar operator<(const Scalar& a, double b);
    Scalar operator>(double a, const Scalar& b);
    Scalar operator<(double a, const Scalar& b);
    Scalar min(const Scalar& a, const Scalar& b);
    Scalar max(const Scalar& a, const Scalar& b);
    Scalar abs(const Scalar& a);
    Scalar sign(const Scalar& a);

    class Vector
    {
    public:
        std::vector<Scalar> v;
        Vector() = default;
        Vector(const std::vector<Scalar>& values) : v(values) {}
        static Vector zero(int32_t n, const Scalar& seed);

        int32_t size() const { return (int32_t)v.size(); }
        std::vector<Scalar>& values() { return v; }
        const std::vector<Scalar>& values() const { return v; }
        Scalar& operator[](int32_t i) { return v[i]; }
        const Scalar& operator[](int32_t i) const { return v[i]; }
        Scalar& operator()(int32_t i) { return v[i]; }
        Vector segment(int32_t begin, int32_t n) const;
        void set_segment(int32_t begin, int32_t n, const Vector& o);

        Scalar dot(const Vector& o) const;
        Vector cross3(const Vector& o) const;
        Scalar squared_norm() const;
        Scalar norm() const;
        Scalar stable_norm(double eps) const;
        Scalar stable_norm(const Scalar& eps) const;
        Vector normalized() const;

        Vector operator+(const Vector& o) const;
        Vector operator-(const Vector& o) const;
        Vector operator*(const Scalar& s) const;
        Vector operator*(double s) const;
        Vector operator/(const Scalar& s) const;
        Vector operator/(double s) const;
        void operator+=(const Vector& o) { *this = *this + o; }
        void operator-=(const Vector& o) { *this = *this - o; }
        void operator*=(const Scalar& s) { *this = *this * s; }
        void operator*=(double s) { *this = *this * s; }
    };
    Vector operator*(double s, const Vector& v);
    Vector operator*(const Scalar& s, const Vector& v);
    Vector operator-(const Vector& v);
    Scalar dot(const Vector& a, const Vector& b);

    class Matrix
    {
    public:
        std::vector<Scalar> m;  // row-major
        int32_t nr = 0, nc = 0;
        Matrix() = default;
        Matrix(const std::vector<Scalar>& values, const std::array<int32_t, 2>& shape);
        static Matrix zero(const std::array<int32_t, 2>& shape, const Scalar& seed);
        static Matrix identity(int32_t n, const Scalar& seed);

        int32_t rows() const { return nr; }
        int32_t cols() const { return nc; }
        std::array<int32_t, 2> shape() const { return { nr, nc }; }
        std::vector<Scalar>& values() { return m; }
        const std::vector<Scalar>& values() const { return m; }
        Scalar& operator()(int32_t i, int32_t j) { return m[i * nc + j]; }
        const Scalar& operator()(int32_t i, int32_t j) const { return m[i * nc + j]; }
        Vector row(int32_t i) const;
        Vector col(int32_t j) const;
        void set_row(int32_t i, const Vector& v);
        void set_col(int32_t j, const Vector& v);
        Matrix block(int32_t r0, int32_t c0, int32_t n_rows, int32_t n_cols) const;
        void set_block(int32_t r0, int32_t c0, int32_t n_rows, int32_t n_cols, const Matrix& b);

        Matrix transpose() const;
        Scalar det() const;
        Matrix inv() const;
        Scalar trace() const;
        Scalar frobenius_norm_sq() const;

        Matrix operator+(const Matrix& o) const;
        Matrix operator-(const Matrix& o) const;
        Matrix operator*(const Matrix& o) const;
        Vector operator*(const Vector& v) const;
        Matrix operator*(const Scalar& s) const;
        Matrix operator*(double s) const;
        Matrix operator/(const Scalar& s) const;
        Matrix operator/(double s) const;
    };
    Matrix operator*(double s, const Matrix& m);
    Matrix operator*(const Scalar& s, const Matrix& m);
    Matrix operator-(const Matrix& m);
    Matrix outer(const Vector& a, const Vector& b);

    class Workspace
    {
    public:
        Workspace();
        Scalar make_scalar();
        Vector make_vector(int32_t n);
        Matrix make_matrix(const std::array<int32_t, 2>& shape);
        std::vector<Scalar> make_scalars(int32_t n);
        std::vector<Vector> make_vectors(int32_t n, int32_t count);
        Scalar get_zero() { return Scalar(g.get(), 0); }
        Scalar get_one() { return Scalar(g.get(), 1); }
        Vector get_zero_vector(int32_t n) { return Vector::zero(n, get_zero()); }
        Matrix get_zero_matrix(const std::array<int32_t, 2>& s) { return Matrix::zero(s, get_zero()); }
        Matrix get_identity_matrix(int32_t n) { return Matrix::identity(n, get_zero()); }
        int32_t get_n_symbols() const { return g->n_symbols(); }
        Graph* graph() { return g.get(); }
        std::shared_ptr<Graph> graph_ptr() { return g; }
    private:
        std::shared_ptr<Graph> g;
    };

    // Differentiation ------------------------------------------------------------
    using DiffCache = std::unordered_map<int64_t, int32_t>;
    Scalar diff(const Scalar& expr, const Scalar& wrt);
    Scalar diff(const Scalar& expr, const Scalar& wrt, DiffCache& cache);
    Vector gradient(const Scalar& expr, const std::vector<Scalar>& wrt);
    Vector gradient(const Scalar& expr, const std::vector<Scalar>& wrt, DiffCache& cache);
    Vector gradient(const Scalar& expr, const Vector& wrt);
    Matrix gradient(const Vector& vec, const std::vector<Scalar>& wrt, bool symmetric);
    Matrix gradient(const Vector& vec, const std::vector<Scalar>& wrt, bool symmetric, DiffCache& cache);
    Matrix hessian(const Scalar& expr, const std::vector<Scalar>& wrt);
    std::vector<Scalar> value_gradient(const Scalar& expr, const std::vector<Scalar>& wrt);
    std::vector<Scalar> value_gradient_hessian(const Scalar& expr, const std::vector<Scalar>& wrt);

    // Flatten helpers (symx/src/symbol/utils.h)
    std::vector<Scalar> collect_scalars(const std::vector<Vector>& vs);
    std::vector<Scalar> collect_scalars(const std::vector<std::vector<Scalar>>& vs);
}
