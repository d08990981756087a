// C++ API tests of the front end on the GPU (built by symx_amd/Makefile as
// symx_amd/lib/api_tests, run by tests/test_cpp_api.py). They exercise the user-facing
// SymX surface the way the reference's own integration tests do:
//   stable_neo_hookean   tests/test_global_potential.cpp:30-182 (FD check of the compiled
//                        derivatives, DoF manipulation, assembled E/grad/Hessian against
//                        a per-element CompiledInLoop sum)
//   conditional          tests/test_global_potential.cpp:184-271 (two quadratic
//                        potentials, one conditional, against the closed form)
//   polynomial_newton    tests/test_newton_method.cpp:10-49
//   compiled_ops         tests/test_compilation.cpp:16-94 (compiled vs host evaluation)
//   add_dof_overloads    GlobalPotential.h:55-61 (array, dynamic vector, row-major matrix
//                        and scalar DoF containers)
//   loop_triangle_area   tests/test_loops.cpp:58-233 (CompiledInLoop against closed-form
//                        triangle areas: summation, activity mask, element counts)
//   empty_potential      a potential with no elements beside a populated one (same solve)
// Each case prints "PASS <name>" or "FAIL <name>: <reason>"; the exit status is the
// number of failed cases.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include <unistd.h>

#include "solver.h"

using namespace symx;

namespace
{
    struct Failure
    {
        std::string what;
    };
    void require(bool ok, const std::string& what)
    {
        if (!ok) throw Failure{ what };
    }
    double norm(const std::vector<double>& a)
    {
        double s = 0.0;
        for (double v : a) s += v * v;
        return std::sqrt(s);
    }
    // Eigen's isApprox: |a - b| <= prec * min(|a|, |b|)
    bool is_approx(const std::vector<double>& a, const std::vector<double>& b, double prec)
    {
        if (a.size() != b.size()) return false;
        std::vector<double> d(a.size());
        for (size_t i = 0; i < a.size(); i++) d[i] = a[i] - b[i];
        return norm(d) <= prec * std::min(norm(a), norm(b));
    }
    bool approx(double a, double b, double tol) { return std::abs(a - b) <= tol * std::max(1.0, std::max(std::abs(a), std::abs(b))); }

    spContext quiet_context()
    {
        spContext c = Context::create();
        c->n_threads = 1;
        c->output->set_enabled(false);
        return c;
    }

    void stable_neo_hookean()
    {
        const int n_elements = 100;
        std::vector<std::array<double, 3>> x1 = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 1, 1, 1 } };
        const std::vector<std::array<double, 3>> X = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
        const std::array<double, 2> lame = { 1e4, 0.3 };
        const std::vector<std::array<int32_t, 5>> tets(n_elements, { 0, 0, 1, 2, 3 });
        // rest Jacobian (columns X_i - X_0) and volume
        std::vector<std::array<double, 9>> J(1);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) J[0][3 * r + c] = X[c + 1][r] - X[0][r];
        const double det = J[0][0] * (J[0][4] * J[0][8] - J[0][5] * J[0][7]) - J[0][1] * (J[0][3] * J[0][8] - J[0][5] * J[0][6]) +
                           J[0][2] * (J[0][3] * J[0][7] - J[0][4] * J[0][6]);
        const std::vector<double> volume = { std::abs(det) / 6.0 };

        auto energy = [&](MappedWorkspace<double>& mws, Element& tet) {
            Vector lm = mws.make_vector(lame);
            Scalar vol = mws.make_scalar(volume, tet[0]);
            Matrix Jm = mws.make_matrix(J, { 3, 3 }, tet[0]);
            std::vector<Vector> x = mws.make_vectors(x1, tet.slice(1, 5));
            Scalar E = lm[0];
            Scalar nu = lm[1];
            Scalar mu = E / (2.0 * (1.0 + nu));
            Scalar lambda = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
            Scalar mu_ = 4.0 / 3.0 * mu;
            Scalar lambda_ = lambda + 5.0 / 6.0 * mu;
            Scalar alpha = 1.0 + mu_ / lambda_ - mu_ / (4.0 * lambda_);
            Matrix j({ x[1][0] - x[0][0], x[2][0] - x[0][0], x[3][0] - x[0][0],
                       x[1][1] - x[0][1], x[2][1] - x[0][1], x[3][1] - x[0][1],
                       x[1][2] - x[0][2], x[2][2] - x[0][2], x[3][2] - x[0][2] }, { 3, 3 });
            Matrix F = j * Jm.inv();
            Scalar detF = F.det();
            Scalar Ic = F.frobenius_norm_sq();
            Scalar psi = 0.5 * mu_ * (Ic - 3.0) + 0.5 * lambda_ * (detF - alpha).powN(2) - 0.5 * mu_ * log(Ic + 1.0);
            return vol * psi;
        };

        spGlobalPotential G = GlobalPotential::create();
        G->add_potential("stable_neo_hookean_tet4", tets, [&](MappedWorkspace<double>& mws, Element& tet) { return energy(mws, tet); });
        G->add_dof(x1);
        spContext context = quiet_context();
        SecondOrderCompiledGlobal compiled(G, context);
        const double fd = compiled.test_derivatives_with_FD(1e-7);
        require(fd < 0.002, "FD check " + std::to_string(fd));

        // DoF manipulation
        std::vector<double> u(12, 1.0), du(12, 1.0);
        G->get_dofs(u.data());
        for (int i = 0; i < 12; i++) u[i] += du[i];
        require(x1[0][0] == 0.0, "x before set_dofs");
        G->set_dofs(u.data());
        require(x1[0][0] == 1.0, "x after set_dofs");
        for (double& v : du) v = -v;
        G->apply_dof_increment(du.data());
        require(x1[0][0] == 0.0, "x after apply_dof_increment");

        double P = 0.0;
        std::vector<double> grad;
        spElementHessians eh = compiled.evaluate_P__dP_du__local_d2P_du2(P, grad);
        const std::vector<double> H = eh->assemble_global(1, 12)->to_dense();

        // the same expression through a per-element loop
        auto [mws, conn] = MappedWorkspace<double>::create(tets);
        Element tet = conn;
        Scalar e = energy(*mws, tet);
        std::vector<Scalar> dofs;
        for (const DataMap& m : mws->get_maps())
            if (m.id() == reinterpret_cast<std::uintptr_t>(&x1))
                for (int k = 0; k < 3; k++) dofs.push_back(Scalar(mws->workspace().graph(), mws->workspace().graph()->symbol_node(m.first_symbol_idx + k)));
        std::vector<Scalar> vals = value_gradient_hessian(e, dofs);
        CompiledInLoop<double> loop(mws, vals, "snh_loop");
        double sum_E = 0.0;
        std::vector<double> sum_g(12, 0.0), sum_H(144, 0.0);
        loop.run(1, [&](View<double> s, int32_t, int32_t, View<int32_t>) {
            sum_E += s[0];
            for (int i = 0; i < 12; i++) sum_g[i] += s[1 + i];
            for (int i = 0; i < 144; i++) sum_H[i] += s[13 + i];
        });
        require(approx(P, sum_E, 1e-8), "energy vs loop sum");
        require(is_approx(grad, sum_g, 1e-8), "gradient vs loop sum");
        require(is_approx(H, sum_H, 1e-8), "Hessian vs loop sum");
    }

    void conditional()
    {
        const int n = 100;
        const double k_ = 1.0;
        std::mt19937_64 rng(7);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        std::vector<std::array<double, 3>> x(n);
        std::vector<std::array<int32_t, 1>> conn(n);
        for (int i = 0; i < n; i++) {
            x[i] = { U(rng), U(rng), U(rng) };
            conn[i] = { i };
        }
        spGlobalPotential G = GlobalPotential::create();
        G->add_potential("attract_to_zero", conn, [&](MappedWorkspace<double>& mws, Element& el) {
            Vector xv = mws.make_vector(x, el[0]);
            Scalar k = mws.make_scalar(k_);
            return 0.5 * k * xv.squared_norm();
        });
        G->add_potential("attract_to_zero_if_z_plus", conn, [&](MappedWorkspace<double>& mws, Element& el) -> std::pair<Scalar, Scalar> {
            Vector xv = mws.make_vector(x, el[0]);
            Scalar k = mws.make_scalar(k_);
            return { 0.5 * k * xv.squared_norm(), xv[2] > 0.0 };
        });
        G->add_dof(x);
        spContext context = quiet_context();
        SecondOrderCompiledGlobal compiled(G, context);
        const double fd = compiled.test_derivatives_with_FD(1e-5);
        require(fd < 1e-5, "FD check " + std::to_string(fd));
        double P_s = 0.0;
        std::vector<double> g_s;
        const std::vector<double> H_s = compiled.evaluate_P__dP_du__local_d2P_du2(P_s, g_s)->assemble_global(1, 3 * n)->to_dense();
        double P = 0.0;
        std::vector<double> g(3 * n, 0.0), H((size_t)9 * n * n, 0.0);
        for (int i = 0; i < n; i++) {
            const int reps = x[i][2] > 0.0 ? 2 : 1;
            for (int r = 0; r < reps; r++) {
                P += 0.5 * k_ * (x[i][0] * x[i][0] + x[i][1] * x[i][1] + x[i][2] * x[i][2]);
                for (int d = 0; d < 3; d++) {
                    g[3 * i + d] += k_ * x[i][d];
                    H[(size_t)(3 * i + d) * 3 * n + 3 * i + d] += k_;
                }
            }
        }
        require(approx(P_s, P, 1e-8), "energy");
        require(is_approx(g_s, g, 1e-8), "gradient");
        require(is_approx(H_s, H, 1e-8), "Hessian");
    }

    void polynomial_newton()
    {
        std::vector<std::array<int32_t, 1>> conn = { { 0 } };
        std::vector<std::array<double, 3>> target = { { 1.0, 2.0, 3.0 } };
        std::vector<std::array<double, 3>> x = { { -1.0, -1.0, -1.0 } };
        spGlobalPotential G = GlobalPotential::create();
        G->add_potential("Polynomial", conn, [&](MappedWorkspace<double>& mws, Element& el) {
            Vector xv = mws.make_vector(x, el[0]);
            Vector t = mws.make_vector(target, el[0]);
            return (xv - t).squared_norm().powN(2);
        });
        G->add_dof(x);
        spContext context = quiet_context();
        NewtonsMethod newton(G, context);
        const SolverReturn ret = newton.solve();
        require(ret == SolverReturn::Successful, std::string("solver returned ") + to_string(ret));
        for (int d = 0; d < 3; d++) require(std::abs(x[0][d] - target[0][d]) <= 1e-1, "x far from the target");
    }

    // A row-major matrix of doubles with data() and size(), as EigenMatrixRM<double> offers.
    struct RowMajorMatrix
    {
        std::vector<double> v;
        double* data() { return v.data(); }
        const double* data() const { return v.data(); }
        int32_t size() const { return (int32_t)v.size(); }
    };

    // The DoF containers of GlobalPotential::add_dof (GlobalPotential.h:55-61): an array of
    // 3-vectors, a dynamic vector and a row-major matrix, in one problem; Newton on
    // 0.5 sum |x_i - t_i|^2 (+ a quartic term) reaches the targets in every container, and the
    // flattened DoF vector concatenates the sets in the order they were added.
    void add_dof_overloads()
    {
        std::vector<std::array<double, 3>> a = { { 0.0, 0.0, 0.0 }, { 1.0, 1.0, 1.0 } };
        std::vector<double> b = { 2.0, 2.0, 2.0 };
        RowMajorMatrix c{ { 3.0, 3.0, 3.0, 4.0, 4.0, 4.0 } };
        const std::vector<std::array<double, 3>> ta = { { 0.5, -0.25, 1.0 }, { 2.0, 0.0, -1.0 } }, tb = { { -3.0, 0.125, 7.0 } },
                                                 tc = { { 1.0, 2.0, 3.0 }, { -1.0, -2.0, -3.0 } };
        spGlobalPotential G = GlobalPotential::create();
        std::vector<std::array<int32_t, 1>> ca = { { 0 }, { 1 } }, cb = { { 0 } }, cc = { { 0 }, { 1 } };
        auto pot = [](const Vector& x, const Vector& t) { return 0.5 * (x - t).squared_norm() + (x - t).squared_norm().powN(2); };
        G->add_potential("a", ca, [&](MappedWorkspace<double>& mws, Element& el) { return pot(mws.make_vector(a, el[0]), mws.make_vector(ta, el[0])); });
        G->add_potential("b", cb, [&](MappedWorkspace<double>& mws, Element& el) { return pot(mws.make_vector(b, 3, el[0]), mws.make_vector(tb, el[0])); });
        G->add_potential("c", cc, [&](MappedWorkspace<double>& mws, Element& el) { return pot(mws.make_vector(c, 3, el[0]), mws.make_vector(tc, el[0])); });
        G->add_dof(a);
        G->add_dof(b);
        G->add_dof(c);
        require(G->get_total_n_dofs() == 15, "15 DoFs in three sets");
        const std::vector<int32_t> off = G->get_dofs_offsets();
        require(off.size() >= 3 && off[0] == 0 && off[1] == 6 && off[2] == 9, "DoF set offsets");
        std::vector<double> u(15);
        G->get_dofs(u.data());
        require(u[6] == 2.0 && u[9] == 3.0 && u[14] == 4.0, "flattened DoFs in set order");
        spContext context = quiet_context();
        NewtonsMethod newton(G, context);
        newton.settings.residual_tolerance_abs = 1e-10;
        const SolverReturn ret = newton.solve();
        require(ret == SolverReturn::Successful, std::string("solver returned ") + to_string(ret));
        for (int i = 0; i < 2; i++)
            for (int d = 0; d < 3; d++) {
                require(std::abs(a[i][d] - ta[i][d]) <= 1e-6, "array DoF set");
                require(std::abs(c.v[3 * i + d] - tc[i][d]) <= 1e-6, "row-major matrix DoF set");
            }
        for (int d = 0; d < 3; d++) require(std::abs(b[d] - tb[0][d]) <= 1e-6, "dynamic vector DoF set");
        // a scalar DoF registers one DoF
        double s = 1.0;
        spGlobalPotential G2 = GlobalPotential::create();
        G2->add_dof(s);
        require(G2->get_total_n_dofs() == 1, "scalar DoF set");
    }

    // A potential added with an empty connectivity (no elements) contributes nothing: the solve
    // takes the same Newton and CG iterations to the same DoFs as without it.
    void empty_potential()
    {
        const std::vector<std::array<double, 3>> t = { { 0.5, -0.25, 1.0 }, { 2.0, 0.0, -1.0 }, { -3.0, 0.125, 7.0 } };
        auto solve = [&](bool with_empty, std::vector<std::array<double, 3>>& x) {
            x = { { 0.0, 0.0, 0.0 }, { 1.0, 1.0, 1.0 }, { 2.0, 2.0, 2.0 } };
            spGlobalPotential G = GlobalPotential::create();
            std::vector<std::array<int32_t, 2>> pairs = { { 0, 1 }, { 1, 2 } }, none;
            auto energy = [&](MappedWorkspace<double>& mws, Element& el) {
                Vector a = mws.make_vector(x, el[0]), b = mws.make_vector(x, el[1]);
                Vector ta = mws.make_vector(t, el[0]), tb = mws.make_vector(t, el[1]);
                return 0.5 * (a - ta).squared_norm() + 0.5 * (b - tb).squared_norm() + ((a - b) - (ta - tb)).squared_norm().powN(2);
            };
            G->add_potential("pairs", pairs, energy);
            if (with_empty) G->add_potential("none", none, energy);
            G->add_dof(x);
            spContext context = quiet_context();
            NewtonsMethod newton(G, context);
            newton.settings.residual_tolerance_abs = 1e-10;
            const SolverReturn ret = newton.solve();
            require(ret == SolverReturn::Successful, std::string("solver returned ") + to_string(ret));
            return std::make_pair(newton.get_last_solve_stats().newton_iterations, newton.get_last_solve_stats().cg_iterations);
        };
        std::vector<std::array<double, 3>> x0, x1;
        const auto it0 = solve(false, x0), it1 = solve(true, x1);
        require(it0 == it1, "same Newton and CG iteration counts with an empty potential");
        for (int i = 0; i < 3; i++)
            for (int d = 0; d < 3; d++) {
                require(x0[i][d] == x1[i][d], "same DoFs with an empty potential");
                require(std::abs(x1[i][d] - t[i][d]) <= 1e-6, "DoFs reach the targets");
            }
    }

    void compiled_ops()
    {
        Workspace ws;
        Scalar a = ws.make_scalar();
        Scalar b = ws.make_scalar();
        std::vector<Scalar> exprs = { a + b, a - b, a * b, a / b, sqrt(a), exp(b), log(a), sin(a), cos(b), tan(a), asin(a / 4.0), acos(b / 4.0),
                                      atan(a), a.powN(3), a.powF(2.5), branch(a - b, a, b), branch(b - a, a, b), log(a - 2.0) };
        Compiled<double> c(exprs, "api_ops", get_codegen_dir());
        require(c.is_valid() && c.get_n_inputs() == 2 && c.get_n_outputs() == (int)exprs.size(), "compiled function shape");
        auto expect_at = [](double av, double bv) {
            return std::vector<double>{ av + bv, av - bv, av * bv, av / bv, std::sqrt(av), std::exp(bv), std::log(av), std::sin(av),
                                        std::cos(bv), std::tan(av), std::asin(av / 4.0), std::acos(bv / 4.0), std::atan(av),
                                        av * av * av, std::pow(av, 2.5), av - bv > 0.0 ? av : bv, bv - av > 0.0 ? av : bv, std::log(av - 2.0) };
        };
        auto check_out = [&](const double* got, double av, double bv, double tol, const std::string& what) {
            const std::vector<double> e = expect_at(av, bv);
            for (size_t i = 0; i < exprs.size(); i++) {
                const bool ok = std::isinf(e[i]) || std::isnan(e[i]) ? (std::isinf(got[i]) && got[i] < 0) : approx(got[i], e[i], tol);
                require(ok, what + " output " + std::to_string(i) + ": " + std::to_string(got[i]) + " vs " + std::to_string(e[i]));
            }
        };
        // single calls on the host, per thread buffers (Compiled.h:44-91)
        c.set_n_threads(3);
        const double in[3][2] = { { 1.3, 0.7 }, { 1.9, 1.1 }, { 0.6, 1.7 } };
        for (int t = 0; t < 3; t++) {
            c.set(a, in[t][0], t);
            c.set(b, in[t][1], t);
        }
        for (int t = 0; t < 3; t++) {
            View<double> got = c.run(t);
            require(got.size() == (int)exprs.size(), "output view size");
            check_out(got.data(), in[t][0], in[t][1], 1e-14, "host thread " + std::to_string(t));
        }
        // the same expressions as a GPU batch: equal to the host values to rounding
        std::vector<double> bin = { 1.3, 0.7, 1.9, 1.1, 0.6, 1.7 }, bout(3 * exprs.size());
        c.run_batch(bin.data(), 3, bout.data());
        for (int t = 0; t < 3; t++) check_out(bout.data() + t * exprs.size(), in[t][0], in[t][1], 1e-13, "batch item " + std::to_string(t));
        // NaN checks only in check mode (static_methods.h:15-16)
        Scalar s = ws.make_scalar();
        Compiled<double> cn({ sqrt(s) }, "api_nan", get_codegen_dir());
        cn.set(s, -1.0);
        require(std::isnan(cn.run()[0]), "sqrt(-1) without check mode");
        check_mode_ON(true);
        bool threw = false;
        try {
            cn.run();
        }
        catch (const std::runtime_error&) {
            threw = true;
        }
        check_mode_ON(false);
        require(threw, "check mode rejects a NaN output");
        // matrix inputs, row major (Compiled.cpp:92-108), and the cache: a second object with
        // the same cache_id loads the compiled code, another cache_id does not
        Matrix M = ws.make_matrix({ 2, 2 });
        const std::vector<Scalar> mexpr = { M.det(), M.trace() };
        const std::string cid = "api_matrix_v1_" + std::to_string((long long)getpid());
        Compiled<double> cm(mexpr, "api_matrix", get_codegen_dir(), cid);
        require(!cm.was_cached(), "first compilation is not cached");
        const double mv[4] = { 2.0, 3.0, -1.0, 4.0 };
        cm.set(M, mv);
        View<double> r = cm.run();
        require(approx(r[0], 11.0, 1e-14) && approx(r[1], 6.0, 1e-14), "matrix det / trace");
        Compiled<double> cm2(mexpr, "api_matrix", get_codegen_dir(), cid);
        require(cm2.was_cached() && cm2.get_compilation_info().was_cached, "same cache_id loads the cached code");
        Compiled<double> cm3;
        require(!cm3.load_if_cached("api_matrix", get_codegen_dir(), cid + "_other"), "another cache_id does not load");
        require(cm3.load_if_cached("api_matrix", get_codegen_dir(), cid), "load_if_cached by name and cache_id");
        cm3.set(M, mv);
        require(approx(cm3.run()[0], 11.0, 1e-14), "loaded code evaluates");
    }

    // A 5 x 5 quad grid over [-0.5, 0.5]^2 (z = 0), two triangles per quad (as
    // generate_triangle_grid, mesh_generation.cpp:329-406, builds it up to the order of the
    // triangles, which the checks below do not depend on).
    void triangle_grid(std::vector<std::array<double, 3>>& V, std::vector<std::array<int32_t, 3>>& T)
    {
        const int q = 5, n = q + 1;
        const double h = 1.0 / q;
        V.clear();
        T.clear();
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) V.push_back({ -0.5 + i * h, -0.5 + j * h, 0.0 });
        for (int i = 0; i < q; i++)
            for (int j = 0; j < q; j++) {
                const int a = n * i + j, b = n * (i + 1) + j, c = n * i + j + 1, d = n * (i + 1) + j + 1;
                T.push_back({ a, b, d });
                T.push_back({ a, d, c });
            }
    }

    void loop_triangle_area_case(bool use_conditional, int n_summations, int num_elements)
    {
        std::vector<std::array<double, 3>> V;
        std::vector<std::array<int32_t, 3>> T;
        triangle_grid(V, T);
        if (num_elements > 0 && num_elements < (int)T.size()) T.resize(num_elements);
        auto [mws, conn] = MappedWorkspace<double>::create(T);
        std::vector<Vector> x = mws->make_vectors(V, conn);
        const Scalar area = 0.5 * (x[1] - x[0]).cross3(x[2] - x[0]).norm();
        std::vector<Scalar> exprs;
        if (n_summations > 1) {
            std::vector<std::array<double, 1>> it(n_summations, std::array<double, 1>{ 1.0 });
            exprs.push_back(mws->add_for_each(it, [&](Vector& v) { return area * v[0]; }));
        }
        else {
            exprs.push_back(area);
        }
        CompiledInLoop<double> loop(mws, exprs, "triangle_area_" + std::to_string(n_summations), get_codegen_dir());
        const size_t ne = T.size();
        std::vector<double> ref(ne);
        std::vector<uint8_t> active(ne, 1);
        double ref_sum = 0.0;
        for (size_t e = 0; e < ne; e++) {
            const auto &A = V[T[e][0]], &B = V[T[e][1]], &C = V[T[e][2]];
            const double ux = B[0] - A[0], uy = B[1] - A[1], vx = C[0] - A[0], vy = C[1] - A[1];
            ref[e] = 0.5 * std::abs(ux * vy - uy * vx) * n_summations;
            const double cx = (A[0] + B[0] + C[0]) / 3.0;
            if (use_conditional) active[e] = cx > 0.0 ? 1 : 0;
            if (active[e]) ref_sum += ref[e];
        }
        std::vector<int> calls(ne, 0);
        int32_t last = -1;
        double sum = 0.0;
        auto cb = [&](View<double> sol, int32_t e, int32_t, View<int32_t> c) {
            require(e >= 0 && (size_t)e < ne, "element index out of range");
            require(e > last, "callbacks out of element order");
            last = e;
            require(active[e] != 0, "callback for an inactive element " + std::to_string(e));
            for (int k = 0; k < 3; k++) require(c[k] == T[e][k], "connectivity view of element " + std::to_string(e));
            require(approx(sol[0], ref[e], 1e-13), "area of element " + std::to_string(e) + ": " + std::to_string(sol[0]) + " vs " + std::to_string(ref[e]));
            calls[e]++;
            sum += sol[0];
        };
        if (use_conditional) loop.run(4, cb, active);
        else loop.run(4, cb);
        for (size_t e = 0; e < ne; e++)
            require(calls[e] == (active[e] ? 1 : 0), "element " + std::to_string(e) + " called " + std::to_string(calls[e]) + " times");
        require(approx(sum, ref_sum, 1e-13), "area sum " + std::to_string(sum) + " vs " + std::to_string(ref_sum));
    }

    void loop_triangle_area()
    {
        for (bool cond : { false, true })
            for (int ns : { 1, 3 })
                for (int num : { -1, 1, 4, 8, 9 }) loop_triangle_area_case(cond, ns, num);
    }
    Scalar nh_density(const Matrix& F, const Scalar& l, const Scalar& m)
    {
        Matrix C = F.transpose() * F;
        Scalar Ic = C.trace();
        Scalar logdetF = log(F.det());
        return 0.5 * m * (Ic - 3) - m * logdetF + 0.5 * l * logdetF.powN(2);
    }

    // The element body of oracle/ref_driver.cpp (run_element, examples/NeoHookean.cpp:12-117),
    // compiled against this front end instead of the reference: prints E, gradient, Hessian.
    int run_element(int argc, char** argv)
    {
        if (argc < 2 + 2 + 24) return 2;
        const double lambda = std::atof(argv[2]), mu = std::atof(argv[3]);
        std::vector<double> xin(12), Xin(12);
        for (int i = 0; i < 12; i++) xin[i] = std::atof(argv[4 + i]);
        for (int i = 0; i < 12; i++) Xin[i] = std::atof(argv[16 + i]);
        auto jac_tet4 = [](Workspace& ws, const std::vector<Vector>& xh) {
            Matrix J = ws.get_identity_matrix(3);
            J.set_col(0, xh[1] - xh[0]);
            J.set_col(1, xh[2] - xh[0]);
            J.set_col(2, xh[3] - xh[0]);
            return J;
        };
        Workspace ws;
        std::vector<Vector> x = ws.make_vectors(3, 4);
        std::vector<Vector> X = ws.make_vectors(3, 4);
        Scalar l = ws.make_scalar();
        Scalar m = ws.make_scalar();
        Matrix Jx = jac_tet4(ws, x);
        Matrix JX = jac_tet4(ws, X);
        Matrix F = Jx * JX.inv();
        Scalar E = JX.det() / 6.0 * nh_density(F, l, m);
        DiffCache dc;
        std::vector<Scalar> dofs = collect_scalars(x);
        Vector g = gradient(E, dofs, dc);
        Matrix h = gradient(g, dofs, true, dc);
        std::vector<Scalar> outs = collect_scalars({ { E }, g.values(), h.values() });
        Compiled<double> c(outs, "ref_nh_element", symx::get_codegen_dir());
        for (int v = 0; v < 4; v++) {
            for (int d = 0; d < 3; d++) {
                c.set(x[v][d], xin[3 * v + d]);
                c.set(X[v][d], Xin[3 * v + d]);
            }
        }
        c.set(l, lambda);
        c.set(m, mu);
        View<double> r = c.run();
        for (int i = 0; i < 157; i++) std::printf("%.17g\n", r[i]);
        return 0;
    }
}

int main(int argc, char** argv)
{
    if (argc > 1 && std::string(argv[1]) == "element") return run_element(argc, argv);
    const std::vector<std::pair<std::string, std::function<void()>>> cases = {
        { "stable_neo_hookean", stable_neo_hookean },
        { "conditional", conditional },
        { "polynomial_newton", polynomial_newton },
        { "compiled_ops", compiled_ops },
        { "loop_triangle_area", loop_triangle_area },
        { "add_dof_overloads", add_dof_overloads },
        { "empty_potential", empty_potential },
    };
    int failed = 0;
    for (const auto& [name, f] : cases) {
        try {
            f();
            std::printf("PASS %s\n", name.c_str());
        }
        catch (const Failure& e) {
            std::printf("FAIL %s: %s\n", name.c_str(), e.what.c_str());
            failed++;
        }
        catch (const std::exception& e) {
            std::printf("FAIL %s: exception %s\n", name.c_str(), e.what());
            failed++;
        }
        std::fflush(stdout);
    }
    return failed;
}
