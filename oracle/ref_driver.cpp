// Driver of the REAL reference (SymX built from /root/reference by oracle/Makefile
// into oracle/_ref/). TEST INFRASTRUCTURE ONLY: it produces reference outputs for
// parity checks and the CPU baseline; nothing in symx_amd/ links or runs it.
//
//   ref_driver element <lambda> <mu> <x0..x11> <X0..X11>
//       Prints E, gradient[12], Hessian[144] of the Neo-Hookean Tet4 element built
//       like examples/NeoHookean.cpp:12-117 (Compiled<double>).
//   ref_driver solve <in.bin> <out.bin> [llt]
//       Builds the Tet4 block problem (Neo-Hookean via fem_integrator, backward
//       Euler inertia, penalty boundary) with GlobalPotential and solves one step
//       with NewtonsMethod (examples/dynamic_elasticity_with_contact.cpp:85-125,
//       examples/rubber_extrusion.cpp:68-107). Input/output layouts are in
//       tests/ref_io.py.
//   ref_driver adjoint <in.bin> <out.bin>
//       Adjoint dL/dlog_mu of the quasistatic cantilever at given positions, composed
//       from the reference's Hessian assembly, DirectLLT solver and Compiled<double>.
//   ref_driver hex_element <in.bin> <out.bin>
//       E, gradient and Hessian of one Neo-Hookean Hex8 / Hex27 / Tet10 element per case
//       (8, 27 or 10 nodes), built with
//       the reference's fem_integrator / fem_jacobian (fem_integrators.cpp:41-133) and
//       evaluated by SecondOrderCompiledGlobal (unprojected element Hessian).
//   ref_driver solve_hex <in.bin> <out.bin>
//       One implicit step of a hexahedral or Tet10 block (Neo-Hookean with per-element mu via
//       fem_integrator, backward Euler inertia, penalty boundary) with NewtonsMethod.
#include <symx>
#include <Eigen/SparseCholesky>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <vector>

using namespace symx;

static Scalar nh_density(const Matrix& F, const Scalar& l, const Scalar& m)
{
    Matrix C = F.transpose() * F;
    Scalar Ic = C.trace();
    Scalar logdetF = log(F.det());
    return 0.5 * m * (Ic - 3) - m * logdetF + 0.5 * l * logdetF.powN(2);
}

static int run_element(int argc, char** argv)
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

template <typename T>
static void rd(std::ifstream& f, T* p, size_t n) { f.read(reinterpret_cast<char*>(p), sizeof(T) * n); }
template <typename T>
static void wr(std::ofstream& f, const T* p, size_t n) { f.write(reinterpret_cast<const char*>(p), sizeof(T) * n); }

static int run_solve(int argc, char** argv)
{
    if (argc < 4) return 2;
    std::ifstream f(argv[2], std::ios::binary);
    int32_t hdr[8];
    rd(f, hdr, 8);
    const int nv = hdr[0], ne = hdr[1], nbc = hdr[2], has_inertia = hdr[3], threads = hdr[4], projection_mode = hdr[5], mirror = hdr[6],
              max_iter = hdr[7];
    double prm[12];
    rd(f, prm, 12);
    const double lambda = prm[0], mu = prm[1], dt = prm[2], k = prm[3];
    const Eigen::Vector3d gravity(prm[4], prm[5], prm[6]);
    struct Data
    {
        std::vector<Eigen::Vector3d> x, X, x0, v;
        std::vector<double> mass;
        std::vector<int32_t> tets, all;
        LabelledConnectivity<2> bc{ { "constraint_idx", "vertex_idx" } };
        std::vector<Eigen::Vector3d> bc_target;
        double lambda, mu, dt, k;
        Eigen::Vector3d gravity;
    } d;
    d.lambda = lambda;
    d.mu = mu;
    d.dt = dt;
    d.k = k;
    d.gravity = gravity;
    d.tets.resize((size_t)ne * 4);
    rd(f, d.tets.data(), d.tets.size());
    auto rdv = [&](std::vector<Eigen::Vector3d>& a, int n) {
        std::vector<double> t((size_t)n * 3);
        rd(f, t.data(), t.size());
        a.resize(n);
        for (int i = 0; i < n; i++) a[i] = Eigen::Vector3d(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    };
    rdv(d.X, nv);
    rdv(d.x, nv);
    rdv(d.x0, nv);
    rdv(d.v, nv);
    d.mass.resize(nv);
    rd(f, d.mass.data(), nv);
    std::vector<int32_t> bcv(nbc);
    rd(f, bcv.data(), nbc);
    rdv(d.bc_target, nbc);
    for (int i = 0; i < nbc; i++) d.bc.numbered_push_back({ bcv[i] });
    d.all.resize(nv);
    for (int i = 0; i < nv; i++) d.all[i] = i;

    spGlobalPotential G = GlobalPotential::create();
    G->add_potential("neohookean_tet4", d.tets, 4, [&](MappedWorkspace<double>& mws, Element& elem) {
        std::vector<Vector> x = mws.make_vectors(d.x, elem);
        std::vector<Vector> X = mws.make_vectors(d.X, elem);
        Scalar l = mws.make_scalar(d.lambda);
        Scalar m = mws.make_scalar(d.mu);
        return fem_integrator(mws, FEM_Element::Tet4, [&](Scalar& w, Vector xi) {
            Matrix Jx = fem_jacobian(FEM_Element::Tet4, x, xi);
            Matrix JX = fem_jacobian(FEM_Element::Tet4, X, xi);
            Matrix F = Jx * JX.inv();
            return nh_density(F, l, m) * JX.det() * w;
        });
    });
    if (has_inertia) {
        G->add_potential("be_inertia", d.all, 1, [&](MappedWorkspace<double>& mws, Element& elem) {
            Vector x1 = mws.make_vector(d.x, elem[0]);
            Vector x0 = mws.make_vector(d.x0, elem[0]);
            Vector v0 = mws.make_vector(d.v, elem[0]);
            Scalar m = mws.make_scalar(d.mass, elem[0]);
            Vector g = mws.make_vector(d.gravity);
            Scalar dts = mws.make_scalar(d.dt);
            Vector xhat = x0 + dts * v0 + dts * dts * g;
            Vector dev = x1 - xhat;
            return 0.5 * m * dev.squared_norm() / dts.powN(2);
        });
    }
    if (nbc > 0) {
        G->add_potential("boundary_conditions", d.bc, [&](MappedWorkspace<double>& mws, Element& conn) {
            Vector x = mws.make_vector(d.x, conn["vertex_idx"]);
            Vector target = mws.make_vector(d.bc_target, conn["constraint_idx"]);
            Scalar ks = mws.make_scalar(d.k);
            return 0.5 * ks * (x - target).squared_norm();
        });
    }
    G->add_dof(d.x);
    spContext context = Context::create();
    context->compilation_directory = symx::get_codegen_dir();
    context->n_threads = threads;
    NewtonsMethod newton(G, context);
    const ProjectionToPD modes[4] = { ProjectionToPD::Newton, ProjectionToPD::ProjectedNewton, ProjectionToPD::ProjectOnDemand,
                                      ProjectionToPD::Progressive };
    newton.settings.projection_mode = modes[projection_mode >= 0 && projection_mode < 4 ? projection_mode : 1];
    newton.settings.project_to_pd_use_mirroring = mirror != 0;
    newton.settings.max_iterations = max_iter;
    newton.settings.residual_tolerance_abs = prm[7];
    newton.settings.residual_tolerance_rel = prm[8];
    newton.settings.cg_rel_tolerance = prm[9];
    newton.settings.cg_abs_tolerance = prm[10];
    newton.settings.projection_eps = prm[11];
    if (argc > 4 && std::string(argv[4]) == "llt") newton.settings.linear_solver = LinearSolver::DirectLLT;
    const auto t0 = std::chrono::steady_clock::now();
    SolverReturn ret = newton.solve();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto& st = newton.get_last_solve_stats();
    // The final energy is evaluated by the caller (a second compiled global in the same
    // process would re-dlopen the same shared-object names).
    const double E = std::numeric_limits<double>::quiet_NaN();
    std::ofstream o(argv[3], std::ios::binary);
    int32_t ohdr[4] = { (int32_t)ret, st.newton_iterations, st.cg_iterations, st.ls_bt_iterations };
    wr(o, ohdr, 4);
    double od[2] = { E, secs };
    wr(o, od, 2);
    std::vector<double> xo((size_t)nv * 3);
    for (int i = 0; i < nv; i++)
        for (int c = 0; c < 3; c++) xo[3 * (size_t)i + c] = d.x[i][c];
    wr(o, xo.data(), xo.size());
    return 0;
}

// Adjoint of a quasistatic equilibrium composed from the reference's public pieces
// (SymX itself has no adjoint): SecondOrderCompiledGlobal element Hessians assembled
// without projection (NewtonsMethod.cpp:102,269), Eigen SimplicialLLT as in its
// DirectLLT branch (NewtonsMethod.cpp:395-418), and the mixed partial
// d(lambda . dE_e/dx)/dlog_mu_e from a Compiled<double> of the same element energy.
static int run_adjoint(int argc, char** argv)
{
    if (argc < 4) return 2;
    std::ifstream f(argv[2], std::ios::binary);
    int32_t hdr[4];
    rd(f, hdr, 4);
    const int nv = hdr[0], ne = hdr[1], nbc = hdr[2], threads = hdr[3];
    double prm[5];
    rd(f, prm, 5);
    struct Data
    {
        std::vector<Eigen::Vector3d> x, X, bc_target;
        std::vector<double> mass, log_mu, dLdx;
        std::vector<int32_t> tets_enum, all;
        LabelledConnectivity<2> bc{ { "constraint_idx", "vertex_idx" } };
        double ratio, k;
        Eigen::Vector3d gravity;
    } d;
    d.ratio = prm[0];
    d.k = prm[1];
    d.gravity = Eigen::Vector3d(prm[2], prm[3], prm[4]);
    std::vector<int32_t> tets((size_t)ne * 4);
    rd(f, tets.data(), tets.size());
    auto rdv = [&](std::vector<Eigen::Vector3d>& a, int n) {
        std::vector<double> t((size_t)n * 3);
        rd(f, t.data(), t.size());
        a.resize(n);
        for (int i = 0; i < n; i++) a[i] = Eigen::Vector3d(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    };
    rdv(d.X, nv);
    rdv(d.x, nv);
    d.mass.resize(nv);
    rd(f, d.mass.data(), nv);
    std::vector<int32_t> bcv(nbc);
    rd(f, bcv.data(), nbc);
    rdv(d.bc_target, nbc);
    d.log_mu.resize(ne);
    rd(f, d.log_mu.data(), ne);
    d.dLdx.resize((size_t)nv * 3);
    rd(f, d.dLdx.data(), d.dLdx.size());
    for (int i = 0; i < nbc; i++) d.bc.numbered_push_back({ bcv[i] });
    d.all.resize(nv);
    for (int i = 0; i < nv; i++) d.all[i] = i;
    d.tets_enum.resize((size_t)ne * 5);
    for (int e = 0; e < ne; e++) {
        d.tets_enum[5 * (size_t)e] = e;
        for (int a = 0; a < 4; a++) d.tets_enum[5 * (size_t)e + 1 + a] = tets[4 * (size_t)e + a];
    }

    // Element energy of tet e as a function of its 4 positions and log(mu).
    auto element_energy = [&](std::vector<Vector>& x, std::vector<Vector>& X, const Scalar& logmu, auto&& integrate) {
        Scalar m = exp(logmu);
        Scalar l = d.ratio * m;
        return integrate([&](Scalar& w, Vector xi) {
            Matrix Jx = fem_jacobian(FEM_Element::Tet4, x, xi);
            Matrix JX = fem_jacobian(FEM_Element::Tet4, X, xi);
            Matrix F = Jx * JX.inv();
            return nh_density(F, l, m) * JX.det() * w;
        });
    };

    spGlobalPotential G = GlobalPotential::create();
    G->add_potential("neohookean_tet4_logmu", d.tets_enum, 5, [&](MappedWorkspace<double>& mws, Element& elem) {
        std::vector<Vector> x = mws.make_vectors(d.x, elem.slice(1, 5));
        std::vector<Vector> X = mws.make_vectors(d.X, elem.slice(1, 5));
        Scalar logmu = mws.make_scalar(d.log_mu, elem[0]);
        return element_energy(x, X, logmu, [&](auto&& summand) { return fem_integrator(mws, FEM_Element::Tet4, summand); });
    });
    G->add_potential("gravity", d.all, 1, [&](MappedWorkspace<double>& mws, Element& elem) {
        Vector x = mws.make_vector(d.x, elem[0]);
        Scalar m = mws.make_scalar(d.mass, elem[0]);
        Vector g = mws.make_vector(d.gravity);
        return -m * g.dot(x);
    });
    if (nbc > 0) {
        G->add_potential("boundary_conditions", d.bc, [&](MappedWorkspace<double>& mws, Element& conn) {
            Vector x = mws.make_vector(d.x, conn["vertex_idx"]);
            Vector target = mws.make_vector(d.bc_target, conn["constraint_idx"]);
            Scalar ks = mws.make_scalar(d.k);
            return 0.5 * ks * (x - target).squared_norm();
        });
    }
    G->add_dof(d.x);
    spContext context = Context::create();
    context->compilation_directory = symx::get_codegen_dir();
    context->n_threads = threads;
    SecondOrderCompiledGlobal compiled(G, context);
    double E = 0.0;
    Eigen::VectorXd grad;
    spElementHessians eh = compiled.evaluate_P__dP_du__local_d2P_du2(E, grad);
    const int ndofs = 3 * nv;
    ElementHessians::spBSM H = eh->assemble_global(threads, ndofs);
    std::vector<Eigen::Triplet<double>> triplets;
    H->to_triplets(triplets);
    Eigen::SparseMatrix<double> A(ndofs, ndofs);
    A.setFromTriplets(triplets.begin(), triplets.end());
    Eigen::SimplicialLLT<Eigen::SparseMatrix<double>> llt;
    llt.compute(A);
    int32_t ok = llt.info() == Eigen::Success ? 1 : 0;
    Eigen::VectorXd rhs = Eigen::Map<const Eigen::VectorXd>(d.dLdx.data(), ndofs);
    Eigen::VectorXd lam = Eigen::VectorXd::Zero(ndofs);
    if (ok) {
        lam = llt.solve(rhs);
        ok = llt.info() == Eigen::Success ? 1 : 0;
    }

    // Mixed partial of one element: d(lambda_e . dE_e/dx_e)/dlog_mu_e.
    Workspace ws;
    std::vector<Vector> xs = ws.make_vectors(3, 4);
    std::vector<Vector> Xs = ws.make_vectors(3, 4);
    Scalar logmu = ws.make_scalar();
    std::vector<Vector> ls = ws.make_vectors(3, 4);
    // fem_jacobian differentiates with respect to xi, so the (single) Tet4 quadrature
    // point enters as symbols {w, xi} whose values are set from the rule
    // ({weight, xi0, xi1, xi2}, fem_integrators.cpp:52-58).
    Scalar qw = ws.make_scalar();
    Vector qxi = ws.make_vector(3);
    const std::vector<std::array<double, 4>> rule = get_integration_rule(FEM_Element::Tet4);
    if (rule.size() != 1) return 3;
    Scalar Ee = element_energy(xs, Xs, logmu, [&](auto&& summand) { return summand(qw, qxi); });
    Vector ge = gradient(Ee, collect_scalars(xs));
    std::vector<Scalar> lflat = collect_scalars(ls);
    Scalar lg = ws.get_zero();
    for (int i = 0; i < 12; i++) lg += lflat[i] * ge[i];
    Scalar dp = diff(lg, logmu);
    Compiled<double> cmix({ dp }, "ref_adjoint_mixed_partial", symx::get_codegen_dir());
    std::vector<double> dLdp(ne);
    for (int e = 0; e < ne; e++) {
        for (int a = 0; a < 4; a++) {
            const int v = tets[4 * (size_t)e + a];
            for (int c = 0; c < 3; c++) {
                cmix.set(xs[a][c], d.x[v][c]);
                cmix.set(Xs[a][c], d.X[v][c]);
                cmix.set(ls[a][c], lam[3 * v + c]);
            }
        }
        cmix.set(logmu, d.log_mu[e]);
        cmix.set(qw, rule[0][0]);
        for (int i = 0; i < 3; i++) cmix.set(qxi[i], rule[0][1 + i]);
        View<double> r = cmix.run();
        dLdp[e] = -r[0];
    }
    std::ofstream o(argv[3], std::ios::binary);
    wr(o, &ok, 1);
    wr(o, &E, 1);
    wr(o, dLdp.data(), dLdp.size());
    wr(o, lam.data(), (size_t)ndofs);
    return 0;
}

// Neo-Hookean hexahedron (nn = 8 or 27 nodes) or quadratic tetrahedron (nn = 10) through the
// reference's own isoparametric integrator: sum over the element's quadrature points of
// psi(F) det(J_X) w.
static Scalar nh_hex_energy(MappedWorkspace<double>& mws, int nn, std::vector<Vector>& x, std::vector<Vector>& X, const Scalar& l,
                            const Scalar& m)
{
    const FEM_Element el = nn == 27 ? FEM_Element::Hex27 : (nn == 10 ? FEM_Element::Tet10 : FEM_Element::Hex8);
    return fem_integrator(mws, el, [&](Scalar& w, Vector& xi) {
        Matrix Jx = fem_jacobian(el, x, xi);
        Matrix JX = fem_jacobian(el, X, xi);
        Matrix F = Jx * JX.inv();
        return nh_density(F, l, m) * JX.det() * w;
    });
}

static int run_hex_element(int argc, char** argv)
{
    if (argc < 4) return 2;
    std::ifstream f(argv[2], std::ios::binary);
    int32_t hdr[2];
    rd(f, hdr, 2);
    const int nn = hdr[0], ncase = hdr[1];
    if (nn != 8 && nn != 10 && nn != 27) return 3;
    const int n = 3 * nn;
    struct
    {
        std::vector<Eigen::Vector3d> x, X;
        std::vector<int32_t> conn;
        double lambda = 0.0, mu = 0.0;
    } d;
    d.x.resize(nn);
    d.X.resize(nn);
    d.conn.resize(nn);
    for (int a = 0; a < nn; a++) d.conn[a] = a;
    spGlobalPotential G = GlobalPotential::create();
    G->add_potential("nh_hex_element", d.conn, nn, [&](MappedWorkspace<double>& mws, Element& elem) {
        std::vector<Vector> x = mws.make_vectors(d.x, elem);
        std::vector<Vector> X = mws.make_vectors(d.X, elem);
        Scalar l = mws.make_scalar(d.lambda);
        Scalar m = mws.make_scalar(d.mu);
        return nh_hex_energy(mws, nn, x, X, l, m);
    });
    G->add_dof(d.x);
    spContext context = Context::create();
    context->compilation_directory = symx::get_codegen_dir();
    context->n_threads = 1;
    SecondOrderCompiledGlobal compiled(G, context);
    std::ofstream o(argv[3], std::ios::binary);
    for (int c = 0; c < ncase; c++) {
        double prm[2];
        rd(f, prm, 2);
        d.lambda = prm[0];
        d.mu = prm[1];
        std::vector<double> t((size_t)2 * n);
        rd(f, t.data(), t.size());
        for (int a = 0; a < nn; a++) {
            d.x[a] = Eigen::Vector3d(t[3 * a], t[3 * a + 1], t[3 * a + 2]);
            d.X[a] = Eigen::Vector3d(t[n + 3 * a], t[n + 3 * a + 1], t[n + 3 * a + 2]);
        }
        double E = 0.0;
        Eigen::VectorXd grad;
        spElementHessians eh = compiled.evaluate_P__dP_du__local_d2P_du2(E, grad);
        ElementHessians::spBSM H = eh->assemble_global(1, n);
        std::vector<Eigen::Triplet<double>> triplets;
        H->to_triplets(triplets);
        std::vector<double> dense((size_t)n * n, 0.0);
        for (const auto& tr : triplets) dense[(size_t)tr.row() * n + tr.col()] += tr.value();
        wr(o, &E, 1);
        wr(o, grad.data(), (size_t)n);
        wr(o, dense.data(), dense.size());
    }
    return 0;
}

static int run_solve_hex(int argc, char** argv)
{
    if (argc < 4) return 2;
    std::ifstream f(argv[2], std::ios::binary);
    int32_t hdr[9];
    rd(f, hdr, 9);
    const int nv = hdr[0], ne = hdr[1], nbc = hdr[2], has_inertia = hdr[3], threads = hdr[4], projection_mode = hdr[5], mirror = hdr[6],
              max_iter = hdr[7], nn = hdr[8];
    if (nn != 8 && nn != 10 && nn != 27) return 3;
    double prm[11];
    rd(f, prm, 11);
    struct Data
    {
        std::vector<Eigen::Vector3d> x, X, x0, v, bc_target;
        std::vector<double> mass, mu;
        std::vector<int32_t> elems, all;
        LabelledConnectivity<2> bc{ { "constraint_idx", "vertex_idx" } };
        double ratio, dt, k;
        Eigen::Vector3d gravity;
    } d;
    d.ratio = prm[0];
    d.dt = prm[1];
    d.k = prm[2];
    d.gravity = Eigen::Vector3d(prm[3], prm[4], prm[5]);
    d.elems.resize((size_t)ne * nn);
    rd(f, d.elems.data(), d.elems.size());
    auto rdv = [&](std::vector<Eigen::Vector3d>& a, int n) {
        std::vector<double> t((size_t)n * 3);
        rd(f, t.data(), t.size());
        a.resize(n);
        for (int i = 0; i < n; i++) a[i] = Eigen::Vector3d(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    };
    rdv(d.X, nv);
    rdv(d.x, nv);
    rdv(d.x0, nv);
    rdv(d.v, nv);
    d.mass.resize(nv);
    rd(f, d.mass.data(), nv);
    d.mu.resize(ne);
    rd(f, d.mu.data(), ne);
    std::vector<int32_t> bcv(nbc);
    rd(f, bcv.data(), nbc);
    rdv(d.bc_target, nbc);
    for (int i = 0; i < nbc; i++) d.bc.numbered_push_back({ bcv[i] });
    d.all.resize(nv);
    for (int i = 0; i < nv; i++) d.all[i] = i;
    // connectivity with the element index first, so the per-element mu is gathered like a vertex value
    std::vector<int32_t> elems_enum((size_t)ne * (nn + 1));
    for (int e = 0; e < ne; e++) {
        elems_enum[(size_t)e * (nn + 1)] = e;
        for (int a = 0; a < nn; a++) elems_enum[(size_t)e * (nn + 1) + 1 + a] = d.elems[(size_t)e * nn + a];
    }

    spGlobalPotential G = GlobalPotential::create();
    G->add_potential("neohookean_hex", elems_enum, nn + 1, [&](MappedWorkspace<double>& mws, Element& elem) {
        std::vector<Vector> x = mws.make_vectors(d.x, elem.slice(1, nn + 1));
        std::vector<Vector> X = mws.make_vectors(d.X, elem.slice(1, nn + 1));
        Scalar m = mws.make_scalar(d.mu, elem[0]);
        Scalar l = mws.make_scalar(d.ratio) * m;
        return nh_hex_energy(mws, nn, x, X, l, m);
    });
    if (has_inertia) {
        G->add_potential("be_inertia", d.all, 1, [&](MappedWorkspace<double>& mws, Element& elem) {
            Vector x1 = mws.make_vector(d.x, elem[0]);
            Vector x0 = mws.make_vector(d.x0, elem[0]);
            Vector v0 = mws.make_vector(d.v, elem[0]);
            Scalar m = mws.make_scalar(d.mass, elem[0]);
            Vector g = mws.make_vector(d.gravity);
            Scalar dts = mws.make_scalar(d.dt);
            Vector xhat = x0 + dts * v0 + dts * dts * g;
            Vector dev = x1 - xhat;
            return 0.5 * m * dev.squared_norm() / dts.powN(2);
        });
    }
    if (nbc > 0) {
        G->add_potential("boundary_conditions", d.bc, [&](MappedWorkspace<double>& mws, Element& conn) {
            Vector x = mws.make_vector(d.x, conn["vertex_idx"]);
            Vector target = mws.make_vector(d.bc_target, conn["constraint_idx"]);
            Scalar ks = mws.make_scalar(d.k);
            return 0.5 * ks * (x - target).squared_norm();
        });
    }
    G->add_dof(d.x);
    spContext context = Context::create();
    context->compilation_directory = symx::get_codegen_dir();
    context->n_threads = threads;
    NewtonsMethod newton(G, context);
    const ProjectionToPD modes[4] = { ProjectionToPD::Newton, ProjectionToPD::ProjectedNewton, ProjectionToPD::ProjectOnDemand,
                                      ProjectionToPD::Progressive };
    newton.settings.projection_mode = modes[projection_mode >= 0 && projection_mode < 4 ? projection_mode : 1];
    newton.settings.project_to_pd_use_mirroring = mirror != 0;
    newton.settings.max_iterations = max_iter;
    newton.settings.residual_tolerance_abs = prm[6];
    newton.settings.residual_tolerance_rel = prm[7];
    newton.settings.cg_rel_tolerance = prm[8];
    newton.settings.cg_abs_tolerance = prm[9];
    newton.settings.projection_eps = prm[10];
    const auto t0 = std::chrono::steady_clock::now();
    SolverReturn ret = newton.solve();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto& st = newton.get_last_solve_stats();
    const double E = std::numeric_limits<double>::quiet_NaN();
    std::ofstream o(argv[3], std::ios::binary);
    int32_t ohdr[4] = { (int32_t)ret, st.newton_iterations, st.cg_iterations, st.ls_bt_iterations };
    wr(o, ohdr, 4);
    double od[2] = { E, secs };
    wr(o, od, 2);
    std::vector<double> xo((size_t)nv * 3);
    for (int i = 0; i < nv; i++)
        for (int c = 0; c < 3; c++) xo[3 * (size_t)i + c] = d.x[i][c];
    wr(o, xo.data(), xo.size());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "element") return run_element(argc, argv);
    if (mode == "solve") return run_solve(argc, argv);
    if (mode == "adjoint") return run_adjoint(argc, argv);
    if (mode == "hex_element") return run_hex_element(argc, argv);
    if (mode == "solve_hex") return run_solve_hex(argc, argv);
    return 2;
}
