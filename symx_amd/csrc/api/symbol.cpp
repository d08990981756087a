#include "symbol.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace symx
{
    const char* op_name(Op op)
    {
        static const char* names[] = { "Zero", "One", "Const", "Symbol", "Add", "Sub", "Mul", "Recip", "PowN", "PowF",
            "Sqrt", "Ln", "Log10", "Exp", "Sin", "Cos", "Tan", "ASin", "ACos", "ATan", "Branch" };
        return names[(int)op];
    }

    size_t Graph::KeyHash::operator()(const std::array<int64_t, 4>& k) const
    {
        uint64_t h = 1469598103934665603ull;
        for (int64_t v : k) {
            h ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        }
        return (size_t)h;
    }

    Graph::Graph()
    {
        Node z; z.op = Op::Zero;
        Node o; o.op = Op::One; o.value = 1.0;
        nodes.push_back(z);
        nodes.push_back(o);
    }

    int32_t Graph::find_or_add(const Node& n)
    {
        int64_t bits = 0;
        if (n.op == Op::Const) std::memcpy(&bits, &n.value, 8);
        std::array<int64_t, 4> key = { (int64_t)n.op, (int64_t)n.a, (int64_t)n.b, n.op == Op::Const ? bits : (int64_t)n.c };
        auto it = cse.find(key);
        if (it != cse.end()) return it->second;
        const int32_t id = (int32_t)nodes.size();
        nodes.push_back(n);
        cse.emplace(key, id);
        return id;
    }

    int32_t Graph::constant(double v)
    {
        // Snap to the identities like the reference (Expressions.cpp:41-60).
        const double tol = 10.0 * std::numeric_limits<double>::epsilon();
        if (std::abs(v) < tol) return 0;
        if (std::abs(v - 1.0) < tol) return 1;
        Node n; n.op = Op::Const; n.value = v;
        return find_or_add(n);
    }

    int32_t Graph::symbol()
    {
        Node n; n.op = Op::Symbol; n.a = (int32_t)symbol_nodes.size();
        const int32_t id = (int32_t)nodes.size();
        nodes.push_back(n);  // symbols are never merged
        symbol_nodes.push_back(id);
        symbol_values.push_back(0.0);
        return id;
    }

    static bool is_const_node(const Node& n) { return n.op == Op::Zero || n.op == Op::One || n.op == Op::Const; }
    static double const_value(const Node& n) { return n.op == Op::Zero ? 0.0 : (n.op == Op::One ? 1.0 : n.value); }

    static double apply_unary(Op op, double x, int32_t ib)
    {
        switch (op) {
        case Op::Recip: return 1.0 / x;
        case Op::PowN: return std::pow(x, (double)ib);
        case Op::Sqrt: return std::sqrt(x);
        case Op::Ln: return (x <= 0.0) ? -std::numeric_limits<double>::infinity() : std::log(x);
        case Op::Log10: return (x <= 0.0) ? -std::numeric_limits<double>::infinity() : std::log10(x);
        case Op::Exp: return std::exp(x);
        case Op::Sin: return std::sin(x);
        case Op::Cos: return std::cos(x);
        case Op::Tan: return std::tan(x);
        case Op::ASin: return std::asin(x);
        case Op::ACos: return std::acos(x);
        case Op::ATan: return std::atan(x);
        default: throw std::runtime_error("symx: apply_unary on non-unary op");
        }
    }

    int32_t Graph::op(Op op, int32_t a, int32_t b)
    {
        const Node& na = nodes[a];
        switch (op) {
        case Op::Add: {
            if (a == 0) return b;
            if (b == 0) return a;
            const Node& nb = nodes[b];
            if (is_const_node(na) && is_const_node(nb)) return constant(const_value(na) + const_value(nb));
            if (a > b) std::swap(a, b);
            break;
        }
        case Op::Sub: {
            if (b == 0) return a;
            if (a == b) return 0;
            const Node& nb = nodes[b];
            if (is_const_node(na) && is_const_node(nb)) return constant(const_value(na) - const_value(nb));
            break;
        }
        case Op::Mul: {
            if (a == 0 || b == 0) return 0;
            if (a == 1) return b;
            if (b == 1) return a;
            const Node& nb = nodes[b];
            if (is_const_node(na) && is_const_node(nb)) return constant(const_value(na) * const_value(nb));
            if (a > b) std::swap(a, b);
            break;
        }
        case Op::PowN: {
            if (b == 0) return 1;
            if (b == 1) return a;
            if (a == 1) return 1;
            if (a == 0 && b > 0) return 0;
            if (is_const_node(na)) return constant(apply_unary(Op::PowN, const_value(na), b));
            break;
        }
        case Op::PowF: {
            const Node& nb = nodes[b];
            if (is_const_node(na) && is_const_node(nb)) return constant(std::pow(const_value(na), const_value(nb)));
            break;
        }
        case Op::Recip:
            if (a == 1) return 1;
            [[fallthrough]];
        default:
            if (op == Op::Recip || op == Op::Sqrt || op == Op::Ln || op == Op::Log10 || op == Op::Exp || op == Op::Sin ||
                op == Op::Cos || op == Op::Tan || op == Op::ASin || op == Op::ACos || op == Op::ATan) {
                if (is_const_node(na)) return constant(apply_unary(op, const_value(na), 0));
                b = -1;
            }
            else {
                throw std::runtime_error("symx: Graph::op called with a non-operation type");
            }
        }
        Node n; n.op = op; n.a = a; n.b = b;
        return find_or_add(n);
    }

    std::vector<int32_t> Graph::substitute(const std::vector<int32_t>& roots, const std::unordered_map<int32_t, int32_t>& replace)
    {
        const int32_t n0 = size();
        std::vector<uint8_t> mark(n0, 0);
        std::vector<int32_t> stack(roots.begin(), roots.end());
        while (!stack.empty()) {
            const int32_t id = stack.back();
            stack.pop_back();
            if (id < 0 || mark[id]) continue;
            mark[id] = 1;
            if (replace.count(id)) continue;
            const Node& n = nodes[id];
            if (n.op == Op::Symbol || n.op == Op::Zero || n.op == Op::One || n.op == Op::Const) continue;
            stack.push_back(n.a);
            if (n.op == Op::Branch) {
                stack.push_back(n.b);
                stack.push_back(n.c);
            }
            else if (n.op != Op::PowN && n.b >= 0) {
                stack.push_back(n.b);
            }
        }
        std::vector<int32_t> map(n0, -1);
        for (int32_t id = 0; id < n0; id++) {  // operands precede their users
            if (!mark[id]) continue;
            auto it = replace.find(id);
            if (it != replace.end()) {
                map[id] = it->second;
                continue;
            }
            const Node n = nodes[id];  // copy: op() may grow the node array
            switch (n.op) {
            case Op::Zero: case Op::One: case Op::Const: case Op::Symbol: map[id] = id; break;
            case Op::Branch: map[id] = branch(map[n.c], map[n.a], map[n.b]); break;
            case Op::PowN: map[id] = op(Op::PowN, map[n.a], n.b); break;
            default: map[id] = op(n.op, map[n.a], n.b >= 0 ? map[n.b] : -1); break;
            }
        }
        std::vector<int32_t> out;
        for (int32_t r : roots) out.push_back(r < 0 ? r : map[r]);
        return out;
    }

    int32_t Graph::branch(int32_t cond, int32_t pos, int32_t neg)
    {
        if (pos == neg) return pos;
        const Node& nc = nodes[cond];
        if (is_const_node(nc)) return const_value(nc) > 0.0 ? pos : neg;
        Node n; n.op = Op::Branch; n.a = pos; n.b = neg; n.c = cond;
        return find_or_add(n);
    }

    static double eval_node(const Node& n, const std::vector<double>& val, const std::vector<double>& symv)
    {
        switch (n.op) {
        case Op::Zero: return 0.0;
        case Op::One: return 1.0;
        case Op::Const: return n.value;
        case Op::Symbol: return symv[n.a];
        case Op::Add: return val[n.a] + val[n.b];
        case Op::Sub: return val[n.a] - val[n.b];
        case Op::Mul: return val[n.a] * val[n.b];
        case Op::PowF: return std::pow(val[n.a], val[n.b]);
        case Op::Branch: return val[n.c] > 0.0 ? val[n.a] : val[n.b];
        default: return apply_unary(n.op, val[n.a], n.b);
        }
    }

    double Graph::eval(int32_t id) const
    {
        return eval_many({ id })[0];
    }

    std::vector<double> Graph::eval_many(const std::vector<int32_t>& ids) const
    {
        int32_t top = 0;
        for (int32_t id : ids) top = std::max(top, id);
        std::vector<double> val((size_t)top + 1, 0.0);
        for (int32_t i = 0; i <= top; i++) val[i] = eval_node(nodes[i], val, symbol_values);
        std::vector<double> out;
        out.reserve(ids.size());
        for (int32_t id : ids) out.push_back(val[id]);
        return out;
    }

    std::string Graph::canonical(const std::vector<int32_t>& roots) const
    {
        // Mark reachable nodes, then print them in id order with dense renumbering.
        std::vector<uint8_t> mark(nodes.size(), 0);
        std::vector<int32_t> stack(roots.begin(), roots.end());
        while (!stack.empty()) {
            const int32_t id = stack.back(); stack.pop_back();
            if (id < 0 || mark[id]) continue;
            mark[id] = 1;
            const Node& n = nodes[id];
            if (n.op == Op::Symbol || is_const_node(n)) continue;
            stack.push_back(n.a);
            if (n.op != Op::PowN) stack.push_back(n.b);
            if (n.op == Op::Branch) stack.push_back(n.c);
        }
        std::vector<int32_t> dense(nodes.size(), -1);
        std::ostringstream os;
        os.precision(17);
        int32_t k = 0;
        for (size_t i = 0; i < nodes.size(); i++) {
            if (!mark[i]) continue;
            dense[i] = k++;
            const Node& n = nodes[i];
            os << (int)n.op << ':';
            switch (n.op) {
            case Op::Const: { uint64_t bits; std::memcpy(&bits, &n.value, 8); os << std::hex << bits << std::dec; break; }
            case Op::Symbol: os << 's' << n.a; break;
            case Op::Zero: case Op::One: break;
            case Op::PowN: os << dense[n.a] << '^' << n.b; break;
            default:
                os << dense[n.a];
                if (n.b >= 0) os << ',' << dense[n.b];
                if (n.op == Op::Branch) os << '?' << dense[n.c];
            }
            os << ';';
        }
        os << '|';
        for (int32_t r : roots) os << dense[r] << ',';
        return os.str();
    }

    // Scalar ---------------------------------------------------------------------
    static Graph* same_graph(const Scalar& a, const Scalar& b)
    {
        if (a.g != b.g) throw std::runtime_error("symx: scalars from different workspaces combined");
        return a.g;
    }
    bool Scalar::is_constant() const { return is_const_node(g->node(id)); }
    double Scalar::constant_value() const { return const_value(g->node(id)); }

    Scalar Scalar::operator+(const Scalar& o) const { Graph* gg = same_graph(*this, o); return Scalar(gg, gg->op(Op::Add, id, o.id)); }
    Scalar Scalar::operator-(const Scalar& o) const { Graph* gg = same_graph(*this, o); return Scalar(gg, gg->op(Op::Sub, id, o.id)); }
    Scalar Scalar::operator*(const Scalar& o) const { Graph* gg = same_graph(*this, o); return Scalar(gg, gg->op(Op::Mul, id, o.id)); }
    Scalar Scalar::operator/(const Scalar& o) const { return *this * o.inv(); }
    Scalar Scalar::operator+(double v) const { return *this + make_constant(v); }
    Scalar Scalar::operator-(double v) const { return *this - make_constant(v); }
    Scalar Scalar::operator*(double v) const { return *this * make_constant(v); }
    Scalar Scalar::operator/(double v) const { return *this * make_constant(1.0 / v); }
    Scalar Scalar::powN(int32_t n) const { return Scalar(g, g->op(Op::PowN, id, n)); }
    Scalar Scalar::powF(double p) const { return Scalar(g, g->op(Op::PowF, id, g->constant(p))); }
    Scalar Scalar::inv() const { return Scalar(g, g->op(Op::Recip, id)); }
    Scalar Scalar::sqrt() const { return Scalar(g, g->op(Op::Sqrt, id)); }
    Scalar Scalar::log() const { return ln(); }
    Scalar Scalar::ln() const { return Scalar(g, g->op(Op::Ln, id)); }
    Scalar Scalar::log10() const { return Scalar(g, g->op(Op::Log10, id)); }
    Scalar Scalar::exp() const { return Scalar(g, g->op(Op::Exp, id)); }
    Scalar Scalar::sin() const { return Scalar(g, g->op(Op::Sin, id)); }
    Scalar Scalar::cos() const { return Scalar(g, g->op(Op::Cos, id)); }
    Scalar Scalar::tan() const { return Scalar(g, g->op(Op::Tan, id)); }
    Scalar Scalar::asin() const { return Scalar(g, g->op(Op::ASin, id)); }
    Scalar Scalar::acos() const { return Scalar(g, g->op(Op::ACos, id)); }
    Scalar Scalar::atan() const { return Scalar(g, g->op(Op::ATan, id)); }

    Scalar operator+(double v, const Scalar& s) { return s.make_constant(v) + s; }
    Scalar operator-(double v, const Scalar& s) { return s.make_constant(v) - s; }
    Scalar operator*(double v, const Scalar& s) { return s.make_constant(v) * s; }
    Scalar operator/(double v, const Scalar& s) { return s.make_constant(v) * s.inv(); }
    Scalar operator-(const Scalar& s) { return s.make_constant(-1.0) * s; }
    Scalar powN(const Scalar& s, int32_t n) { return s.powN(n); }
    Scalar powF(const Scalar& s, double p) { return s.powF(p); }
    Scalar sqrt(const Scalar& s) { return s.sqrt(); }
    Scalar log(const Scalar& s) { return s.ln(); }
    Scalar ln(const Scalar& s) { return s.ln(); }
    Scalar log10(const Scalar& s) { return s.log10(); }
    Scalar exp(const Scalar& s) { return s.exp(); }
    Scalar sin(const Scalar& s) { return s.sin(); }
    Scalar cos(const Scalar& s) { return s.cos(); }
    Scalar tan(const Scalar& s) { return s.tan(); }
    Scalar asin(const Scalar& s) { return s.asin(); }
    Scalar acos(const Scalar& s) { return s.acos(); }
    Scalar atan(const Scalar& s) { return s.atan(); }

    Scalar branch(const Scalar& c, const Scalar& p, const Scalar& n)
    {
        Graph* g = same_graph(c, p); same_graph(p, n);
        return Scalar(g, g->branch(c.id, p.id, n.id));
    }
    Scalar branch(const Scalar& c, double p, const Scalar& n) { return branch(c, c.make_constant(p), n); }
    Scalar branch(const Scalar& c, const Scalar& p, double n) { return branch(c, p, c.make_constant(n)); }
    Scalar branch(const Scalar& c, double p, double n) { return branch(c, c.make_constant(p), c.make_constant(n)); }
    // Comparisons produce the signed difference; a branch takes its positive side
    // where that value is > 0 (Scalar.cpp:531-538).
    Scalar operator>(const Scalar& a, const Scalar& b) { return a - b; }
    Scalar operator<(const Scalar& a, const Scalar& b) { return b - a; }
    Scalar operator>(const Scalar& a, double b) { return a - a.make_constant(b); }
    Scalar operator<(const Scalar& a, double b) { return a.make_constant(b) - a; }
    Scalar operator>(double a, const Scalar& b) { return b.make_constant(a) - b; }
    Scalar operator<(double a, const Scalar& b) { return b - b.make_constant(a); }
    Scalar min(const Scalar& a, const Scalar& b) { return branch(a > b, b, a); }
    Scalar max(const Scalar& a, const Scalar& b) { return branch(a > b, a, b); }
    Scalar abs(const Scalar& a) { return branch(a > 0.0, a, -a); }
    Scalar sign(const Scalar& a) { return branch(a > 0.0, 1.0, -1.0); }

    // Vector ---------------------------------------------------------------------
    Vector Vector::zero(int32_t n, const Scalar& seed) { return Vector(std::vector<Scalar>((size_t)n, seed.get_zero())); }
    Vector Vector::segment(int32_t begin, int32_t n) const { return Vector(std::vector<Scalar>(v.begin() + begin, v.begin() + begin + n)); }
    void Vector::set_segment(int32_t begin, int32_t n, const Vector& o) { for (int32_t i = 0; i < n; i++) v[begin + i] = o[i]; }
    Scalar Vector::dot(const Vector& o) const
    {
        if (o.size() != size()) throw std::runtime_error("symx: Vector::dot size mismatch");
        Scalar acc = v[0] * o[0];
        for (int32_t i = 1; i < size(); i++) acc = acc + v[i] * o[i];
        return acc;
    }
    Vector Vector::cross3(const Vector& o) const
    {
        return Vector({ v[1] * o[2] - v[2] * o[1], v[2] * o[0] - v[0] * o[2], v[0] * o[1] - v[1] * o[0] });
    }
    Scalar Vector::squared_norm() const { return dot(*this); }
    Scalar Vector::norm() const { return squared_norm().sqrt(); }
    Scalar Vector::stable_norm(const Scalar& eps) const { return (squared_norm() + eps).sqrt() - eps.sqrt(); }
    Scalar Vector::stable_norm(double eps) const { return stable_norm(v[0].make_constant(eps)); }
    Vector Vector::normalized() const { return *this / norm(); }
    Vector Vector::operator+(const Vector& o) const { Vector r(*this); for (int32_t i = 0; i < size(); i++) r.v[i] = v[i] + o[i]; return r; }
    Vector Vector::operator-(const Vector& o) const { Vector r(*this); for (int32_t i = 0; i < size(); i++) r.v[i] = v[i] - o[i]; return r; }
    Vector Vector::operator*(const Scalar& s) const { Vector r(*this); for (auto& x : r.v) x = x * s; return r; }
    Vector Vector::operator*(double s) const { Vector r(*this); for (auto& x : r.v) x = x * s; return r; }
    Vector Vector::operator/(const Scalar& s) const { return *this * s.inv(); }
    Vector Vector::operator/(double s) const { return *this * (1.0 / s); }
    Vector operator*(double s, const Vector& v) { return v * s; }
    Vector operator*(const Scalar& s, const Vector& v) { Vector r(v); for (auto& x : r.v) x = s * x; return r; }
    Vector operator-(const Vector& v) { return v * -1.0; }
    Scalar dot(const Vector& a, const Vector& b) { return a.dot(b); }

    // Matrix ---------------------------------------------------------------------
    Matrix::Matrix(const std::vector<Scalar>& values, const std::array<int32_t, 2>& shape) : m(values), nr(shape[0]), nc(shape[1])
    {
        if ((int32_t)values.size() != nr * nc) throw std::runtime_error("symx: Matrix shape does not match value count");
    }
    Matrix Matrix::zero(const std::array<int32_t, 2>& shape, const Scalar& seed)
    {
        return Matrix(std::vector<Scalar>((size_t)shape[0] * shape[1], seed.get_zero()), shape);
    }
    Matrix Matrix::identity(int32_t n, const Scalar& seed)
    {
        Matrix r = zero({ n, n }, seed);
        for (int32_t i = 0; i < n; i++) r(i, i) = seed.get_one();
        return r;
    }
    Vector Matrix::row(int32_t i) const { Vector r; for (int32_t j = 0; j < nc; j++) r.v.push_back((*this)(i, j)); return r; }
    Vector Matrix::col(int32_t j) const { Vector r; for (int32_t i = 0; i < nr; i++) r.v.push_back((*this)(i, j)); return r; }
    void Matrix::set_row(int32_t i, const Vector& v) { for (int32_t j = 0; j < nc; j++) (*this)(i, j) = v[j]; }
    void Matrix::set_col(int32_t j, const Vector& v) { for (int32_t i = 0; i < nr; i++) (*this)(i, j) = v[i]; }
    Matrix Matrix::block(int32_t r0, int32_t c0, int32_t n_rows, int32_t n_cols) const
    {
        Matrix r = zero({ n_rows, n_cols }, m[0]);
        for (int32_t i = 0; i < n_rows; i++) for (int32_t j = 0; j < n_cols; j++) r(i, j) = (*this)(r0 + i, c0 + j);
        return r;
    }
    void Matrix::set_block(int32_t r0, int32_t c0, int32_t n_rows, int32_t n_cols, const Matrix& b)
    {
        for (int32_t i = 0; i < n_rows; i++) for (int32_t j = 0; j < n_cols; j++) (*this)(r0 + i, c0 + j) = b(i, j);
    }
    Matrix Matrix::transpose() const
    {
        Matrix r = zero({ nc, nr }, m[0]);
        for (int32_t i = 0; i < nr; i++) for (int32_t j = 0; j < nc; j++) r(j, i) = (*this)(i, j);
        return r;
    }
    Scalar Matrix::det() const
    {
        const Matrix& A = *this;
        if (nr == 2 && nc == 2) return A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
        if (nr == 3 && nc == 3) {
            return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1))
                 - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0))
                 + A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
        }
        throw std::runtime_error("symx: Matrix::det supports 2x2 and 3x3 only");
    }
    Matrix Matrix::inv() const
    {
        const Matrix& A = *this;
        if (nr == 2 && nc == 2) {
            const Scalar id = det().inv();
            return Matrix({ A(1, 1) * id, -A(0, 1) * id, -A(1, 0) * id, A(0, 0) * id }, { 2, 2 });
        }
        if (nr == 3 && nc == 3) {
            // Adjugate / determinant.
            const Scalar c00 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
            const Scalar c01 = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
            const Scalar c02 = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
            const Scalar d = A(0, 0) * c00 + A(0, 1) * c01 + A(0, 2) * c02;
            const Scalar id = d.inv();
            Matrix r = zero({ 3, 3 }, m[0]);
            r(0, 0) = c00 * id;
            r(1, 0) = c01 * id;
            r(2, 0) = c02 * id;
            r(0, 1) = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * id;
            r(1, 1) = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * id;
            r(2, 1) = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * id;
            r(0, 2) = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * id;
            r(1, 2) = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * id;
            r(2, 2) = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * id;
            return r;
        }
        throw std::runtime_error("symx: Matrix::inv supports 2x2 and 3x3 only");
    }
    Scalar Matrix::trace() const
    {
        Scalar acc = (*this)(0, 0);
        for (int32_t i = 1; i < std::min(nr, nc); i++) acc = acc + (*this)(i, i);
        return acc;
    }
    Scalar Matrix::frobenius_norm_sq() const
    {
        Scalar acc = m[0] * m[0];
        for (size_t i = 1; i < m.size(); i++) acc = acc + m[i] * m[i];
        return acc;
    }
    Matrix Matrix::operator+(const Matrix& o) const { Matrix r(*this); for (size_t i = 0; i < m.size(); i++) r.m[i] = m[i] + o.m[i]; return r; }
    Matrix Matrix::operator-(const Matrix& o) const { Matrix r(*this); for (size_t i = 0; i < m.size(); i++) r.m[i] = m[i] - o.m[i]; return r; }
    Matrix Matrix::operator*(const Matrix& o) const
    {
        if (nc != o.nr) throw std::runtime_error("symx: Matrix product shape mismatch");
        Matrix r = zero({ nr, o.nc }, m[0]);
        for (int32_t i = 0; i < nr; i++) {
            for (int32_t j = 0; j < o.nc; j++) {
                Scalar acc = (*this)(i, 0) * o(0, j);
                for (int32_t k = 1; k < nc; k++) acc = acc + (*this)(i, k) * o(k, j);
                r(i, j) = acc;
            }
        }
        return r;
    }
    Vector Matrix::operator*(const Vector& v) const
    {
        Vector r;
        for (int32_t i = 0; i < nr; i++) {
            Scalar acc = (*this)(i, 0) * v[0];
            for (int32_t k = 1; k < nc; k++) acc = acc + (*this)(i, k) * v[k];
            r.v.push_back(acc);
        }
        return r;
    }
    Matrix Matrix::operator*(const Scalar& s) const { Matrix r(*this); for (auto& x : r.m) x = x * s; return r; }
    Matrix Matrix::operator*(double s) const { Matrix r(*this); for (auto& x : r.m) x = x * s; return r; }
    Matrix Matrix::operator/(const Scalar& s) const { return *this * s.inv(); }
    Matrix Matrix::operator/(double s) const { return *this * (1.0 / s); }
    Matrix operator*(double s, const Matrix& m) { return m * s; }
    Matrix operator*(const Scalar& s, const Matrix& m) { Matrix r(m); for (auto& x : r.m) x = s * x; return r; }
    Matrix operator-(const Matrix& m) { return m * -1.0; }
    Matrix outer(const Vector& a, const Vector& b)
    {
        Matrix r = Matrix::zero({ a.size(), b.size() }, a[0]);
        for (int32_t i = 0; i < a.size(); i++) for (int32_t j = 0; j < b.size(); j++) r(i, j) = a[i] * b[j];
        return r;
    }

    // Workspace ------------------------------------------------------------------
    Workspace::Workspace() : g(std::make_shared<Graph>()) {}
    Scalar Workspace::make_scalar() { return Scalar(g.get(), g->symbol()); }
    Vector Workspace::make_vector(int32_t n) { Vector r; for (int32_t i = 0; i < n; i++) r.v.push_back(make_scalar()); return r; }
    Matrix Workspace::make_matrix(const std::array<int32_t, 2>& shape)
    {
        std::vector<Scalar> vals;
        for (int32_t i = 0; i < shape[0] * shape[1]; i++) vals.push_back(make_scalar());
        return Matrix(vals, shape);
    }
    std::vector<Scalar> Workspace::make_scalars(int32_t n) { std::vector<Scalar> r; for (int32_t i = 0; i < n; i++) r.push_back(make_scalar()); return r; }
    std::vector<Vector> Workspace::make_vectors(int32_t n, int32_t count) { std::vector<Vector> r; for (int32_t i = 0; i < count; i++) r.push_back(make_vector(n)); return r; }

    // Differentiation ------------------------------------------------------------
    static int32_t d_impl(Graph* g, int32_t id, int32_t wrt_sym, DiffCache& cache);

    static int32_t d_cached(Graph* g, int32_t id, int32_t wrt_sym, DiffCache& cache)
    {
        const int64_t key = ((int64_t)id << 32) | (uint32_t)wrt_sym;
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
        const int32_t r = d_impl(g, id, wrt_sym, cache);
        cache.emplace(key, r);
        return r;
    }

    static int32_t d_impl(Graph* g, int32_t id, int32_t wrt_sym, DiffCache& cache)
    {
        const Node n = g->node(id);
        switch (n.op) {
        case Op::Zero: case Op::One: case Op::Const: return 0;
        case Op::Symbol: return n.a == wrt_sym ? 1 : 0;
        default: break;
        }
        Scalar u(g, n.a);
        Scalar du(g, d_cached(g, n.a, wrt_sym, cache));
        switch (n.op) {
        case Op::Branch: {
            Scalar dv(g, d_cached(g, n.b, wrt_sym, cache));
            return g->branch(n.c, du.id, dv.id);
        }
        case Op::PowN:
            if (du.is_zero()) return 0;
            return ((double)n.b * u.powN(n.b - 1) * du).id;
        case Op::PowF: {
            // d(u^p) = p u^(p-1) du for a constant exponent.
            Scalar p(g, n.b);
            if (!p.is_constant()) throw std::runtime_error("symx: diff of PowF with a non-constant exponent");
            if (du.is_zero()) return 0;
            return (p * u.powF(p.constant_value() - 1.0) * du).id;
        }
        default: break;
        }
        if (du.is_zero() && n.op != Op::Add && n.op != Op::Sub && n.op != Op::Mul) return 0;
        switch (n.op) {
        case Op::Add: { Scalar dv(g, d_cached(g, n.b, wrt_sym, cache)); return (du + dv).id; }
        case Op::Sub: { Scalar dv(g, d_cached(g, n.b, wrt_sym, cache)); return (du - dv).id; }
        case Op::Mul: {
            Scalar v(g, n.b);
            Scalar dv(g, d_cached(g, n.b, wrt_sym, cache));
            return (u * dv + v * du).id;
        }
        case Op::Recip: return (-du / u.powN(2)).id;
        case Op::Sqrt: return (0.5 / sqrt(u) * du).id;
        case Op::Ln: return (du / u).id;
        case Op::Log10: return (du / (u * std::log(10.0))).id;
        case Op::Exp: return (exp(u) * du).id;
        case Op::Sin: return (cos(u) * du).id;
        case Op::Cos: return (-sin(u) * du).id;
        case Op::Tan: return ((1.0 + tan(u).powN(2)) * du).id;
        case Op::ASin: return (du / sqrt(1.0 - u.powN(2))).id;
        case Op::ACos: return (-du / sqrt(1.0 - u.powN(2))).id;
        case Op::ATan: return (du / (1.0 + u.powN(2))).id;
        default: throw std::runtime_error(std::string("symx: diff of unsupported op ") + op_name(n.op));
        }
    }

    Scalar diff(const Scalar& expr, const Scalar& wrt)
    {
        DiffCache c;
        return diff(expr, wrt, c);
    }
    Scalar diff(const Scalar& expr, const Scalar& wrt, DiffCache& cache)
    {
        same_graph(expr, wrt);
        if (!wrt.is_symbol()) throw std::runtime_error("symx: diff with respect to a non-symbol");
        return Scalar(expr.g, d_cached(expr.g, expr.id, wrt.get_symbol_idx(), cache));
    }
    Vector gradient(const Scalar& expr, const std::vector<Scalar>& wrt) { DiffCache c; return gradient(expr, wrt, c); }
    Vector gradient(const Scalar& expr, const std::vector<Scalar>& wrt, DiffCache& cache)
    {
        Vector r;
        for (const Scalar& s : wrt) r.v.push_back(diff(expr, s, cache));
        return r;
    }
    Vector gradient(const Scalar& expr, const Vector& wrt) { return gradient(expr, wrt.values()); }
    Matrix gradient(const Vector& vec, const std::vector<Scalar>& wrt, bool symmetric) { DiffCache c; return gradient(vec, wrt, symmetric, c); }
    Matrix gradient(const Vector& vec, const std::vector<Scalar>& wrt, bool symmetric, DiffCache& cache)
    {
        const int32_t nr = vec.size(), nc = (int32_t)wrt.size();
        Matrix m = Matrix::zero({ nr, nc }, vec[0]);
        for (int32_t i = 0; i < nr; i++) {
            for (int32_t j = 0; j < nc; j++) {
                if (symmetric && j < i) m(i, j) = m(j, i);
                else m(i, j) = diff(vec[i], wrt[j], cache);
            }
        }
        return m;
    }
    Matrix hessian(const Scalar& expr, const std::vector<Scalar>& wrt)
    {
        DiffCache c;
        Vector g = gradient(expr, wrt, c);
        return gradient(g, wrt, true, c);
    }
    std::vector<Scalar> value_gradient(const Scalar& expr, const std::vector<Scalar>& wrt)
    {
        std::vector<Scalar> r{ expr };
        Vector g = gradient(expr, wrt);
        r.insert(r.end(), g.v.begin(), g.v.end());
        return r;
    }
    std::vector<Scalar> value_gradient_hessian(const Scalar& expr, const std::vector<Scalar>& wrt)
    {
        DiffCache c;
        std::vector<Scalar> r{ expr };
        Vector g = gradient(expr, wrt, c);
        Matrix h = gradient(g, wrt, true, c);
        r.insert(r.end(), g.v.begin(), g.v.end());
        r.insert(r.end(), h.m.begin(), h.m.end());
        return r;
    }
    std::vector<Scalar> collect_scalars(const std::vector<Vector>& vs)
    {
        std::vector<Scalar> r;
        for (const auto& v : vs) r.insert(r.end(), v.v.begin(), v.v.end());
        return r;
    }
    std::vector<Scalar> collect_scalars(const std::vector<std::vector<Scalar>>& vs)
    {
        std::vector<Scalar> r;
        for (const auto& v : vs) r.insert(r.end(), v.begin(), v.end());
        return r;
    }
}
