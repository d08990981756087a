// Example problems behind include/symx_scenes.h, written against the C++ front
// end exactly as a SymX user would write them (cf. examples/NeoHookean.cpp,
// examples/rubber_extrusion.cpp:68-107, examples/dynamic_elasticity_with_contact.cpp:67-102).
#include "../../../include/symx_scenes.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../api/solver.h"

using namespace symx;

namespace
{
    thread_local std::string g_err;

    int guard(const std::function<void()>& f)
    {
        try {
            f();
            return 0;
        }
        catch (const std::exception& e) {
            g_err = e.what();
            return 1;
        }
    }

    // SplitMix64 -> uniform double in [0, 1) with 53 random bits.
    struct SplitMix64
    {
        uint64_t s;
        explicit SplitMix64(uint64_t seed) : s(seed) {}
        uint64_t next()
        {
            uint64_t z = (s += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        }
        double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    };

    Scalar neohookean_energy_density(const Matrix& F, const Scalar& l, const Scalar& m)
    {
        // 0.5 mu (I_C - 3) - mu ln J + 0.5 lambda ln^2 J (examples/elasticity_potentials.h:7-13)
        const Matrix C = F.transpose() * F;
        const Scalar Ic = C.trace();
        const Scalar logdetF = log(F.det());
        return 0.5 * m * (Ic - 3.0) - m * logdetF + 0.5 * l * logdetF.powN(2);
    }

    Scalar stable_neohookean_energy_density(const Matrix& F, const Scalar& youngs, const Scalar& poisson)
    {
        // Smith et al. stable Neo-Hookean (examples/elasticity_potentials.h:15-30)
        const Scalar mu = youngs / (2.0 * (1.0 + poisson));
        const Scalar lambda = (youngs * poisson) / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
        const Scalar mu_ = 4.0 / 3.0 * mu;
        const Scalar lambda_ = lambda + 5.0 / 6.0 * mu;
        const Scalar alpha = 1.0 + mu_ / lambda_ - mu_ / (4.0 * lambda_);
        const Scalar detF = F.det();
        const Scalar Ic = F.frobenius_norm_sq();
        return 0.5 * mu_ * (Ic - 3.0) + 0.5 * lambda_ * (detF - alpha).powN(2) - 0.5 * mu_ * log(Ic + 1.0);
    }

    // Elastic Tet4 potential through fem_integrator, as the reference examples write it.
    Scalar tet4_elastic(MappedWorkspace<double>& mws, const std::vector<Vector>& x, const std::vector<Vector>& X, const Scalar& a,
                        const Scalar& b, int material)
    {
        return fem_integrator(mws, FEM_Element::Tet4, [&](Scalar& w, Vector& xi) {
            const Matrix Jx = fem_jacobian(FEM_Element::Tet4, x, xi);
            const Matrix JX = fem_jacobian(FEM_Element::Tet4, X, xi);
            const Matrix F = Jx * JX.inv();
            const Scalar psi = material == 0 ? neohookean_energy_density(F, a, b) : stable_neohookean_energy_density(F, a, b);
            return psi * JX.det() * w;
        });
    }
}

struct symx_scene_s
{
    symx_tet4_block_params prm{};
    symx_newton_params np{};
    std::vector<std::array<double, 3>> x, x0, v, X;
    std::vector<double> mass;
    std::vector<int32_t> tets;      // [n][4]
    std::vector<int32_t> tets_enum; // [n][5]: element index, v0..v3
    std::vector<int32_t> all_vertices;
    LabelledConnectivity<2> bc{ { "constraint_idx", "vertex_idx" } };
    std::vector<std::array<double, 3>> bc_target;
    std::vector<double> log_mu;
    std::vector<int32_t> contact;   // surface vertices watched by the ground barrier
    double kappa = 0, dhat = 0;
    symx_sphere_params sph{};
    bool is_sphere = false;
    int32_t npe = 4;                  // nodes per element of `tets`
    std::vector<double> E_elem;       // per-element Young's modulus (hexahedral scenes)
    std::vector<int32_t> elems_enum;  // [n][npe + 1]: element index, nodes
    double lambda = 0, mu = 0, lam_over_mu = 0, dt = 0, k = 0;
    double E_ = 0, nu_ = 0;
    std::array<double, 3> gravity{};
    std::array<double, 3> mg{};
    spContext ctx;
    spGlobalPotential G;
    spNewtonsMethod newton;
};

static void apply_newton_params(NewtonSettings& s, const symx_newton_params& p, OutputSink& out)
{
    out.set_console_verbosity(p.verbosity <= 0 ? Verbosity::None : p.verbosity == 1 ? Verbosity::Summary : p.verbosity == 2 ? Verbosity::Medium : Verbosity::Full);
    s.precision = p.precision == SYMX_HIP_MIXED ? Precision::Mixed : Precision::FP64;
    s.projection_mode = p.projection_mode == 0   ? ProjectionToPD::Newton
                        : p.projection_mode == 2 ? ProjectionToPD::ProjectOnDemand
                        : p.projection_mode == 3 ? ProjectionToPD::Progressive
                                                 : ProjectionToPD::ProjectedNewton;
    s.project_to_pd_use_mirroring = p.use_mirroring != 0;
    s.max_iterations = p.max_iterations;
    s.residual_tolerance_abs = p.residual_tolerance_abs;
    s.residual_tolerance_rel = p.residual_tolerance_rel;
    s.cg_rel_tolerance = p.cg_rel_tolerance;
    s.cg_abs_tolerance = p.cg_abs_tolerance;
    s.cg_max_iterations = p.cg_max_iterations;
    s.projection_eps = p.projection_eps;
    s.step_tolerance = p.step_tolerance;
    s.linear_solver = p.linear_solver == 1 ? LinearSolver::DirectLLT : LinearSolver::BDPCG;
}

// Material, loads and contact of a Tet4 scene (shared by the block and sphere scenes).
struct SceneSpec
{
    double youngs_modulus = 1e5, poissons_ratio = 0.4, density = 1000.0;
    std::array<double, 3> gravity{ 0.0, 0.0, -9.81 };
    double dt = 0.01, bc_stiffness = 1e8;
    int32_t material = 0, log_mu_field = 0;
    bool contact = false;
    double kappa = 0.0, dhat = 0.0;
};

static void add_common_potentials(symx_scene_s* S, const SceneSpec& p);

// Masses, per-element data and potentials of a Tet4 scene whose mesh (X, tets), state
// (x, x0, v) and boundary vertices (bc, bc_target) are set (no device work).
static void finish_scene(symx_scene_s* s, const SceneSpec& p)
{
    const int32_t nv = (int32_t)s->X.size();
    const int32_t ne = (int32_t)(s->tets.size() / 4);
    // lumped masses (dynamic_elasticity_with_contact.cpp:67-79)
    s->mass.assign(nv, 0.0);
    for (int32_t e = 0; e < ne; e++) {
        const int32_t* t = &s->tets[4 * (size_t)e];
        const auto& A = s->X[t[0]];
        double d1[3], d2[3], d3[3];
        for (int k = 0; k < 3; k++) {
            d1[k] = s->X[t[1]][k] - A[k];
            d2[k] = s->X[t[2]][k] - A[k];
            d3[k] = s->X[t[3]][k] - A[k];
        }
        const double det = d1[0] * (d2[1] * d3[2] - d2[2] * d3[1]) - d1[1] * (d2[0] * d3[2] - d2[2] * d3[0]) + d1[2] * (d2[0] * d3[1] - d2[1] * d3[0]);
        const double lumped = std::abs(det) / 6.0 * p.density / 4.0;
        for (int k = 0; k < 4; k++) s->mass[t[k]] += lumped;
    }
    s->all_vertices.resize(nv);
    for (int32_t i = 0; i < nv; i++) s->all_vertices[i] = i;
    s->tets_enum.resize((size_t)ne * 5);
    for (int32_t e = 0; e < ne; e++) {
        s->tets_enum[5 * (size_t)e] = e;
        for (int k = 0; k < 4; k++) s->tets_enum[5 * (size_t)e + 1 + k] = s->tets[4 * (size_t)e + k];
    }
    s->E_ = p.youngs_modulus;
    s->nu_ = p.poissons_ratio;
    s->mu = p.youngs_modulus / (2.0 * (1.0 + p.poissons_ratio));
    s->lambda = p.youngs_modulus * p.poissons_ratio / ((1.0 + p.poissons_ratio) * (1.0 - 2.0 * p.poissons_ratio));
    s->lam_over_mu = s->lambda / s->mu;
    s->log_mu.assign(ne, std::log(s->mu));
    s->dt = p.dt;
    s->k = p.bc_stiffness;
    s->gravity = p.gravity;
    s->kappa = p.kappa;
    s->dhat = p.dhat;

    s->G = GlobalPotential::create();
    symx_scene_s* S = s;
    const int material = p.material;
    if (p.log_mu_field) {
        S->G->add_potential("neohookean_tet4_logmu", S->tets_enum, 5, [S](MappedWorkspace<double>& mws, Element& elem) {
            std::vector<Vector> xs = mws.make_vectors(S->x, elem.slice(1, 5));
            std::vector<Vector> Xs = mws.make_vectors(S->X, elem.slice(1, 5));
            const Scalar logmu = mws.make_scalar(S->log_mu, elem[0]);
            const Scalar ratio = mws.make_scalar(S->lam_over_mu);
            const Scalar m = exp(logmu);
            return tet4_elastic(mws, xs, Xs, ratio * m, m, 0);
        });
    }
    else {
        S->G->add_potential(material == 0 ? "neohookean_tet4" : "stable_neohookean_tet4", S->tets, 4,
                            [S, material](MappedWorkspace<double>& mws, Element& elem) {
                                std::vector<Vector> xs = mws.make_vectors(S->x, elem);
                                std::vector<Vector> Xs = mws.make_vectors(S->X, elem);
                                const Scalar a = material == 0 ? mws.make_scalar(S->lambda) : mws.make_scalar(S->E_);
                                const Scalar b = material == 0 ? mws.make_scalar(S->mu) : mws.make_scalar(S->nu_);
                                return tet4_elastic(mws, xs, Xs, a, b, material);
                            });
    }
    add_common_potentials(S, p);
}

// Inertia or gravity, penalty boundary, ground contact, DoFs and parameters.
static void add_common_potentials(symx_scene_s* S, const SceneSpec& p)
{
    if (p.dt > 0.0) {
        // backward Euler inertia with gravity (dynamic_elasticity_with_contact.cpp:85-102, no damping)
        S->G->add_potential("be_inertia", S->all_vertices, 1, [S](MappedWorkspace<double>& mws, Element& elem) {
            const Vector x1 = mws.make_vector(S->x, elem[0]);
            const Vector x0 = mws.make_vector(S->x0, elem[0]);
            const Vector v0 = mws.make_vector(S->v, elem[0]);
            const Scalar m = mws.make_scalar(S->mass, elem[0]);
            const Vector g = mws.make_vector(S->gravity);
            const Scalar dt = mws.make_scalar(S->dt);
            const Vector xhat = x0 + dt * v0 + dt * dt * g;
            const Vector dev = x1 - xhat;
            return 0.5 * m * dev.squared_norm() / dt.powN(2);
        });
    }
    else if (p.gravity[0] != 0.0 || p.gravity[1] != 0.0 || p.gravity[2] != 0.0) {
        S->G->add_potential("gravity", S->all_vertices, 1, [S](MappedWorkspace<double>& mws, Element& elem) {
            const Vector x1 = mws.make_vector(S->x, elem[0]);
            const Scalar m = mws.make_scalar(S->mass, elem[0]);
            const Vector g = mws.make_vector(S->gravity);
            return -1.0 * m * g.dot(x1);
        });
    }
    if (!S->bc_target.empty()) {
        // penalty Dirichlet condition (rubber_extrusion.cpp:95-107)
        S->G->add_potential("boundary_conditions", S->bc, [S](MappedWorkspace<double>& mws, Element& conn) {
            const Vector x = mws.make_vector(S->x, conn["vertex_idx"]);
            const Vector target = mws.make_vector(S->bc_target, conn["constraint_idx"]);
            const Scalar k = mws.make_scalar(S->k);
            return 0.5 * k * (x - target).squared_norm();
        });
    }
    if (p.contact && !S->contact.empty()) {
        // IPC log-barrier against the ground plane z = 0 for each surface vertex, active
        // below dhat: b(d) = -kappa (d - dhat)^2 ln(d / dhat), +inf for d <= 0. Written as a
        // conditional potential (GlobalPotential.h:20,38), like the ground contact of
        // dynamic_elasticity_with_contact.cpp:127-137.
        S->G->add_potential("ipc_ground_barrier", S->contact, 1, [S](MappedWorkspace<double>& mws, Element& elem) -> std::pair<Scalar, Scalar> {
            const Vector x1 = mws.make_vector(S->x, elem[0]);
            const Scalar kappa = mws.make_scalar(S->kappa);
            const Scalar dhat = mws.make_scalar(S->dhat);
            const Scalar d = x1[2];
            const Scalar b = -1.0 * kappa * (d - dhat).powN(2) * log(d / dhat);
            return { b, dhat > d };
        });
    }
    S->G->add_dof(S->x);
    if (p.log_mu_field) S->G->add_parameter(S->log_mu, "log_mu");
}

// Mesh, data and potentials of the Tet4 block scene (no device work).
static void build_scene(symx_scene_s* s, const symx_tet4_block_params* p, const symx_newton_params* np)
{
    s->prm = *p;
    s->np = *np;
    const FEMMesh mesh = generate_cuboid_Tet4_mesh({ p->size[0], p->size[1], p->size[2] }, { p->cells[0], p->cells[1], p->cells[2] });
    const int32_t nv = (int32_t)mesh.vertices.size();
    s->X = mesh.vertices;
    s->tets = mesh.connectivity;
    s->x = s->X;
    SplitMix64 rng(p->seed);
    std::vector<uint8_t> fixed(nv, 0);
    for (int32_t i = 0; i < nv; i++) {
        if (p->fixed_axis >= 0 && std::abs(s->X[i][p->fixed_axis]) < 1e-12) fixed[i] = 1;
        for (int d = 0; d < 3; d++) {
            const double u = rng.uniform();
            if (!fixed[i]) s->x[i][d] += p->noise * (2.0 * u - 1.0);
        }
    }
    if (p->renumber_seed != 0) {
        // vertex i becomes perm[i]
        std::vector<int32_t> perm(nv);
        for (int32_t i = 0; i < nv; i++) perm[i] = i;
        SplitMix64 shuffle(p->renumber_seed);
        for (int32_t i = nv - 1; i > 0; i--) std::swap(perm[i], perm[(int32_t)(shuffle.next() % (uint64_t)(i + 1))]);
        std::vector<std::array<double, 3>> X(nv), x(nv);
        std::vector<uint8_t> f(nv);
        for (int32_t i = 0; i < nv; i++) {
            X[perm[i]] = s->X[i];
            x[perm[i]] = s->x[i];
            f[perm[i]] = fixed[i];
        }
        s->X = std::move(X);
        s->x = std::move(x);
        fixed = std::move(f);
        for (auto& t : s->tets) t = perm[t];
    }
    s->x0 = s->x;
    s->v.assign(nv, { 0.0, 0.0, 0.0 });
    for (int32_t i = 0; i < nv; i++) {
        if (fixed[i]) {
            s->bc.numbered_push_back({ i });
            s->bc_target.push_back(s->X[i]);
        }
    }
    SceneSpec spec;
    spec.youngs_modulus = p->youngs_modulus;
    spec.poissons_ratio = p->poissons_ratio;
    spec.density = p->density;
    spec.gravity = { p->gravity[0], p->gravity[1], p->gravity[2] };
    spec.dt = p->dt;
    spec.bc_stiffness = p->bc_stiffness;
    spec.material = p->material;
    spec.log_mu_field = p->log_mu_field;
    finish_scene(s, spec);
}

// Jittered-grid tetrahedralisation of a ball: the 6-tet split of a cells^3 grid over the
// bounding cube (mesh_generation.cpp:64-84 order), every grid vertex jittered by
// U(-jitter, jitter) cell widths (SplitMix64 stream, 3 draws per grid vertex in index
// order), tets kept when the centroid of their unjittered vertices lies within the
// radius, vertices renumbered in increasing grid order. Contact vertices: vertices of
// faces that belong to one kept tet, ascending.
static void sphere_mesh(const symx_sphere_params* p, std::vector<std::array<double, 3>>& X, std::vector<int32_t>& tets,
                        std::vector<int32_t>& contact)
{
    if (p->cells < 1 || !(p->radius > 0.0) || !(p->jitter >= 0.0 && p->jitter < 0.25)) throw std::runtime_error("bad sphere parameters");
    const double r = p->radius, h = 2.0 * r / p->cells;
    const FEMMesh grid = generate_cuboid_Tet4_mesh({ 2.0 * r, 2.0 * r, 2.0 * r }, { p->cells, p->cells, p->cells },
                                                   { p->center[0] - r, p->center[1] - r, p->center[2] - r });
    const int32_t ng = (int32_t)grid.vertices.size();
    std::vector<std::array<double, 3>> J = grid.vertices;
    SplitMix64 rng(p->seed);
    for (int32_t i = 0; i < ng; i++)
        for (int d = 0; d < 3; d++) J[i][d] += p->jitter * h * (2.0 * rng.uniform() - 1.0);
    std::vector<int32_t> keep;
    for (int32_t e = 0; e < grid.n_elements; e++) {
        double c[3] = { 0.0, 0.0, 0.0 };
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) c[d] += 0.25 * grid.vertices[grid.connectivity[4 * (size_t)e + a]][d];
        const double dx = c[0] - p->center[0], dy = c[1] - p->center[1], dz = c[2] - p->center[2];
        if (dx * dx + dy * dy + dz * dz <= r * r) keep.push_back(e);
    }
    std::vector<int32_t> id(ng, -1);
    for (int32_t e : keep)
        for (int a = 0; a < 4; a++) id[grid.connectivity[4 * (size_t)e + a]] = 0;
    X.clear();
    for (int32_t i = 0; i < ng; i++) {
        if (id[i] < 0) continue;
        id[i] = (int32_t)X.size();
        X.push_back(J[i]);
    }
    tets.resize(keep.size() * 4);
    for (size_t k = 0; k < keep.size(); k++)
        for (int a = 0; a < 4; a++) tets[4 * k + a] = id[grid.connectivity[4 * (size_t)keep[k] + a]];
    // boundary faces
    std::vector<std::array<int32_t, 3>> faces;
    faces.reserve(keep.size() * 4);
    for (size_t k = 0; k < keep.size(); k++) {
        const int32_t* t = &tets[4 * k];
        const int32_t f[4][3] = { { t[0], t[1], t[2] }, { t[0], t[1], t[3] }, { t[0], t[2], t[3] }, { t[1], t[2], t[3] } };
        for (auto& fc : f) {
            std::array<int32_t, 3> q = { fc[0], fc[1], fc[2] };
            std::sort(q.begin(), q.end());
            faces.push_back(q);
        }
    }
    std::sort(faces.begin(), faces.end());
    std::vector<uint8_t> on(X.size(), 0);
    for (size_t i = 0; i < faces.size();) {
        size_t j = i + 1;
        while (j < faces.size() && faces[j] == faces[i]) j++;
        if (j - i == 1)
            for (int32_t v : faces[i]) on[v] = 1;
        i = j;
    }
    contact.clear();
    for (int32_t v = 0; v < (int32_t)X.size(); v++)
        if (on[v]) contact.push_back(v);
}

static void build_sphere(symx_scene_s* s, const symx_sphere_params* p, const symx_newton_params* np)
{
    s->sph = *p;
    s->np = *np;
    s->is_sphere = true;
    sphere_mesh(p, s->X, s->tets, s->contact);
    s->x = s->X;
    s->x0 = s->X;
    s->v.assign(s->X.size(), { p->velocity[0], p->velocity[1], p->velocity[2] });
    SceneSpec spec;
    spec.youngs_modulus = p->youngs_modulus;
    spec.poissons_ratio = p->poissons_ratio;
    spec.density = p->density;
    spec.gravity = { p->gravity[0], p->gravity[1], p->gravity[2] };
    spec.dt = p->dt;
    spec.material = p->material;
    spec.contact = p->contact_stiffness > 0.0;
    spec.kappa = p->contact_stiffness;
    spec.dhat = p->dhat;
    finish_scene(s, spec);
}

// Hexahedral block: Hex8 or Hex27 cuboid over [0, size] (or, with 10 nodes per element, the
// Tet10 block: six quadratic tetrahedra per cell), per-element Young's modulus
// log-uniform in [youngs_min, youngs_max] (SplitMix64 stream seed ^ 0x9e37..., one draw
// per element), node noise like the Tet4 block (stream `seed`, 3 draws per node, noise in
// node spacings), masses lumped uniformly (rho V_e / nodes per element to each node).
static void build_hex(symx_scene_s* s, const symx_hex_block_params* p, const symx_newton_params* np)
{
    if (p->nodes_per_element != 8 && p->nodes_per_element != 10 && p->nodes_per_element != 27)
        throw std::runtime_error("nodes_per_element must be 8, 10 or 27");
    s->np = *np;
    s->npe = p->nodes_per_element;
    const int npe = s->npe;
    // 10 nodes: the Tet10 block of the same cuboid (six quadratic tetrahedra per cell)
    const FEM_Element type = npe == 27 ? FEM_Element::Hex27 : (npe == 10 ? FEM_Element::Tet10 : FEM_Element::Hex8);
    const FEMMesh mesh = generate_cuboid_mesh(type, { p->size[0], p->size[1], p->size[2] }, { p->cells[0], p->cells[1], p->cells[2] });
    const int32_t nv = (int32_t)mesh.vertices.size(), ne = mesh.n_elements;
    s->X = mesh.vertices;
    s->tets = mesh.connectivity;
    s->x = s->X;
    const int sub = npe == 8 ? 1 : 2;
    const double h[3] = { p->size[0] / (sub * p->cells[0]), p->size[1] / (sub * p->cells[1]), p->size[2] / (sub * p->cells[2]) };
    SplitMix64 rng(p->seed);
    std::vector<uint8_t> fixed(nv, 0);
    for (int32_t i = 0; i < nv; i++) {
        if (p->fixed_axis >= 0 && std::abs(s->X[i][p->fixed_axis]) < 1e-12) fixed[i] = 1;
        for (int d = 0; d < 3; d++) {
            const double u = rng.uniform();
            if (!fixed[i]) s->x[i][d] += p->noise * h[d] * (2.0 * u - 1.0);
        }
    }
    s->x0 = s->x;
    s->v.assign(nv, { 0.0, 0.0, 0.0 });
    for (int32_t i = 0; i < nv; i++)
        if (fixed[i]) {
            s->bc.numbered_push_back({ i });
            s->bc_target.push_back(s->X[i]);
        }
    SplitMix64 rng_e(p->seed ^ 0x9e3779b97f4a7c15ull);
    s->E_elem.resize(ne);
    const double l0 = std::log(p->youngs_min), l1 = std::log(p->youngs_max);
    for (int32_t e = 0; e < ne; e++) s->E_elem[e] = std::exp(l0 + rng_e.uniform() * (l1 - l0));
    const double vol = (p->size[0] / p->cells[0]) * (p->size[1] / p->cells[1]) * (p->size[2] / p->cells[2]) / (npe == 10 ? 6.0 : 1.0);
    s->mass.assign(nv, 0.0);
    for (int32_t e = 0; e < ne; e++)
        for (int a = 0; a < npe; a++) s->mass[s->tets[(size_t)npe * e + a]] += p->density * vol / npe;
    s->all_vertices.resize(nv);
    for (int32_t i = 0; i < nv; i++) s->all_vertices[i] = i;
    s->elems_enum.resize((size_t)ne * (npe + 1));
    for (int32_t e = 0; e < ne; e++) {
        s->elems_enum[(size_t)(npe + 1) * e] = e;
        for (int a = 0; a < npe; a++) s->elems_enum[(size_t)(npe + 1) * e + 1 + a] = s->tets[(size_t)npe * e + a];
    }
    s->nu_ = p->poissons_ratio;
    s->dt = p->dt;
    s->k = p->bc_stiffness;
    s->gravity = { p->gravity[0], p->gravity[1], p->gravity[2] };
    s->G = GlobalPotential::create();
    symx_scene_s* S = s;
    S->G->add_potential(npe == 27 ? "neohookean_hex27" : (npe == 10 ? "neohookean_tet10" : "neohookean_hex8"), S->elems_enum, npe + 1,
                        [S, npe, type](MappedWorkspace<double>& mws, Element& elem) {
                            std::vector<Vector> xs = mws.make_vectors(S->x, elem.slice(1, npe + 1));
                            std::vector<Vector> Xs = mws.make_vectors(S->X, elem.slice(1, npe + 1));
                            const Scalar E = mws.make_scalar(S->E_elem, elem[0]);
                            const Scalar nu = mws.make_scalar(S->nu_);
                            const Scalar mu = E / (2.0 * (1.0 + nu));
                            const Scalar lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
                            return fem_integrator(mws, type, [&](Scalar& w, Vector& xi) {
                                const Matrix Jx = fem_jacobian(type, xs, xi);
                                const Matrix JX = fem_jacobian(type, Xs, xi);
                                const Matrix F = Jx * JX.inv();
                                return neohookean_energy_density(F, lambda, mu) * JX.det() * w;
                            });
                        });
    SceneSpec spec;
    spec.density = p->density;
    spec.gravity = s->gravity;
    spec.dt = p->dt;
    spec.bc_stiffness = p->bc_stiffness;
    add_common_potentials(s, spec);
}

extern "C" {

void symx_hex_block_params_default(symx_hex_block_params* p)
{
    p->cells[0] = p->cells[1] = p->cells[2] = 4;
    p->size[0] = p->size[1] = p->size[2] = 1.0;
    p->nodes_per_element = 27;
    p->youngs_min = 1e4;
    p->youngs_max = 1e6;
    p->poissons_ratio = 0.45;
    p->density = 1000.0;
    p->gravity[0] = 0.0;
    p->gravity[1] = 0.0;
    p->gravity[2] = -9.81;
    p->dt = 0.01;
    p->noise = 0.1;
    p->seed = 3;
    p->bc_stiffness = 1e8;
    p->fixed_axis = 2;
}

int symx_scene_hex_block_create(const symx_hex_block_params* p, const symx_newton_params* np, symx_scene* out)
{
    *out = nullptr;
    auto* s = new symx_scene_s();
    int rc = guard([&]() {
        build_hex(s, p, np);
        s->ctx = Context::create();
        s->newton = NewtonsMethod::create(s->G, s->ctx);
        apply_newton_params(s->newton->settings, *np, *s->ctx->output);
    });
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int symx_scene_nodes_per_element(symx_scene s, int32_t* n)
{
    *n = s->npe;
    return 0;
}

int symx_scene_young_moduli(symx_scene s, double* out)
{
    std::memcpy(out, s->E_elem.data(), s->E_elem.size() * sizeof(double));
    return 0;
}

void symx_newton_params_default(symx_newton_params* p)
{
    p->precision = SYMX_HIP_FP64;
    p->projection_mode = 1;
    p->use_mirroring = 0;
    p->max_iterations = 100;
    p->residual_tolerance_abs = 0.0;
    p->residual_tolerance_rel = 1e-6;
    p->cg_rel_tolerance = 1e-3;
    p->cg_abs_tolerance = 1e-12;
    p->cg_max_iterations = 10000;
    p->projection_eps = 1e-10;
    p->step_tolerance = 0.0;
    p->verbosity = 0;
    p->linear_solver = 0;
}

void symx_tet4_block_params_default(symx_tet4_block_params* p)
{
    p->cells[0] = p->cells[1] = p->cells[2] = 20;
    p->size[0] = p->size[1] = p->size[2] = 1.0;
    p->youngs_modulus = 1e5;
    p->poissons_ratio = 0.4;
    p->density = 1000.0;
    p->gravity[0] = 0.0;
    p->gravity[1] = 0.0;
    p->gravity[2] = -9.81;
    p->dt = 0.01;
    p->noise = 0.01;
    p->seed = 1;
    p->bc_stiffness = 1e8;
    p->fixed_axis = 2;
    p->material = 0;
    p->log_mu_field = 0;
    p->renumber_seed = 0;
}

void symx_sphere_params_default(symx_sphere_params* p)
{
    p->radius = 0.5;
    p->cells = 50;
    p->jitter = 0.1;
    p->seed = 7;
    p->center[0] = 0.0;
    p->center[1] = 0.0;
    p->center[2] = 0.5 + 0.01;
    p->velocity[0] = 0.0;
    p->velocity[1] = 0.0;
    p->velocity[2] = -2.0;
    p->youngs_modulus = 1e5;
    p->poissons_ratio = 0.4;
    p->density = 1000.0;
    p->gravity[0] = 0.0;
    p->gravity[1] = 0.0;
    p->gravity[2] = -9.81;
    p->dt = 1.0 / 200.0;
    p->material = 0;
    p->contact_stiffness = 1e6;
    p->dhat = 1e-3;
    p->step_filter = 0.9;
}

const char* symx_scene_last_error(void) { return g_err.c_str(); }

int symx_scene_tet4_block_create(const symx_tet4_block_params* p, const symx_newton_params* np, symx_scene* out)
{
    *out = nullptr;
    auto* s = new symx_scene_s();
    int rc = guard([&]() {
        build_scene(s, p, np);
        symx_scene_s* S = s;
        S->ctx = Context::create();
        S->newton = NewtonsMethod::create(S->G, S->ctx);
        apply_newton_params(S->newton->settings, *np, *S->ctx->output);
    });
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int symx_scene_sphere_create(const symx_sphere_params* p, const symx_newton_params* np, symx_scene* out)
{
    *out = nullptr;
    auto* s = new symx_scene_s();
    int rc = guard([&]() {
        build_sphere(s, p, np);
        s->ctx = Context::create();
        s->newton = NewtonsMethod::create(s->G, s->ctx);
        apply_newton_params(s->newton->settings, *np, *s->ctx->output);
        if (p->contact_stiffness > 0.0 && p->step_filter > 0.0) {
            symx_hip_problem prob = s->newton->device().problem();
            if (symx_hip_problem_set_plane_filter(prob, s->contact.data(), (int32_t)s->contact.size(), 2, 0.0))
                throw std::runtime_error(symx_hip_last_error(nullptr));
            const double eta = p->step_filter;
            s->newton->callbacks->add_max_allowed_step_device([eta](symx_hip_problem pr) {
                double t = 1.0;
                if (symx_hip_problem_plane_max_step(pr, eta, &t)) throw std::runtime_error(symx_hip_last_error(nullptr));
                return t;
            });
        }
    });
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int symx_scene_contact_vertices(symx_scene s, int32_t* n, int32_t* out)
{
    *n = (int32_t)s->contact.size();
    if (out) std::memcpy(out, s->contact.data(), s->contact.size() * sizeof(int32_t));
    return 0;
}

void symx_scene_destroy(symx_scene s)
{
    if (!s) return;
    s->newton.reset();
    s->G.reset();
    s->ctx.reset();
    delete s;
}

int symx_scene_n_vertices(symx_scene s, int32_t* n)
{
    *n = (int32_t)s->x.size();
    return 0;
}
int symx_scene_n_elements(symx_scene s, int32_t* n)
{
    *n = (int32_t)(s->tets.size() / s->npe);
    return 0;
}
int symx_scene_connectivity(symx_scene s, int32_t* conn)
{
    std::memcpy(conn, s->tets.data(), s->tets.size() * sizeof(int32_t));
    return 0;
}

static std::vector<std::array<double, 3>>* vec3_array(symx_scene s, const std::string& name)
{
    if (name == "x") return &s->x;
    if (name == "x0") return &s->x0;
    if (name == "v") return &s->v;
    if (name == "X") return &s->X;
    return nullptr;
}

int symx_scene_get_array(symx_scene s, const char* name, double* out)
{
    if (std::string(name) == "mass") {
        std::memcpy(out, s->mass.data(), s->mass.size() * sizeof(double));
        return 0;
    }
    auto* a = vec3_array(s, name);
    if (!a) {
        g_err = std::string("unknown array ") + name;
        return 1;
    }
    std::memcpy(out, a->data(), a->size() * 3 * sizeof(double));
    return 0;
}
int symx_scene_set_array(symx_scene s, const char* name, const double* in)
{
    auto* a = vec3_array(s, name);
    if (!a || std::string(name) == "X") {
        g_err = std::string("array cannot be set: ") + name;
        return 1;
    }
    std::memcpy(a->data(), in, a->size() * 3 * sizeof(double));
    return 0;
}
int symx_scene_set_log_mu(symx_scene s, const double* log_mu)
{
    return guard([&]() {
        std::memcpy(s->log_mu.data(), log_mu, s->log_mu.size() * sizeof(double));
        s->newton->device().refresh_data();
    });
}
int symx_scene_set_newton_params(symx_scene s, const symx_newton_params* np)
{
    s->np = *np;
    apply_newton_params(s->newton->settings, *np, *s->ctx->output);
    return 0;
}

int symx_scene_solve(symx_scene s, symx_solve_stats* st)
{
    return guard([&]() {
        s->ctx->logger->clear();
        const auto t0 = std::chrono::steady_clock::now();
        const SolverReturn r = s->newton->solve();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st) {
            const auto& ss = s->newton->get_last_solve_stats();
            const auto& L = *s->ctx->logger;
            st->result = (int32_t)r;
            st->newton_iterations = ss.newton_iterations;
            st->cg_iterations = ss.cg_iterations;
            st->ls_bt_iterations = ss.ls_bt_iterations;
            st->ls_trapezoid_accepts = ss.ls_trapezoid_accepts;
            st->n_refinements = ss.n_refinements;
            st->final_energy = ss.final_energy;
            st->final_residual = ss.final_residual;
            st->first_residual = ss.residuals.empty() ? 0.0 : ss.residuals.front();
            st->seconds = secs;
            st->seconds_eval = L.get_timer_total("evaluate_P__dP_du__local_d2P_du2");
            st->seconds_project = L.get_timer_total("project_to_PD");
            st->seconds_assembly = L.get_timer_total("assembly");
            st->seconds_linear = L.get_timer_total("linear_system_solve");
            st->seconds_energy = L.get_timer_total("evaluate_P");
        }
    });
}

int symx_scene_adjoint(symx_scene s, const double* dL_dx, double* dL_dlog_mu, double rel_tol, int32_t max_iter, symx_adjoint_stats* st)
{
    return guard([&]() {
        if (s->log_mu.empty() || s->newton->device().n_parameters() != (int32_t)s->log_mu.size())
            throw std::runtime_error("scene has no per-element log(mu) parameters");
        const std::vector<double> rhs(dL_dx, dL_dx + s->x.size() * 3);
        std::vector<double> grad;
        const auto a = s->newton->adjoint(rhs, grad, rel_tol, max_iter);
        std::memcpy(dL_dlog_mu, grad.data(), grad.size() * sizeof(double));
        if (st) {
            st->cg_iterations = a.cg_iterations;
            st->n_refinements = a.n_refinements;
            st->converged = a.converged;
            st->found_indefiniteness = a.found_indefiniteness;
            st->cg_error = a.cg_error;
            st->energy = a.energy;
            st->seconds = a.seconds;
        }
    });
}

int symx_scene_step(symx_scene s, symx_solve_stats* st)
{
    int rc = symx_scene_solve(s, st);
    if (rc) return rc;
    if (s->dt > 0.0) {
        // v = (x - x0) / dt, x0 = x (dynamic_elasticity_with_contact.cpp's time loop), over
        // host threads: at 1M vertices one thread takes ~8 ms, more than a Newton iteration
        const size_t n = s->x.size();
        auto update = [s](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; i++) {
                for (int d = 0; d < 3; d++) {
                    s->v[i][d] = (s->x[i][d] - s->x0[i][d]) / s->dt;
                    s->x0[i][d] = s->x[i][d];
                }
            }
        };
        const size_t nt = n < 262144 ? 1 : std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; t++) pool.emplace_back(update, n * t / nt, n * (t + 1) / nt);
        update(0, n / nt);
        for (auto& th : pool) th.join();
    }
    return 0;
}

symx_hip_problem symx_scene_problem(symx_scene s) { return s->newton->device().problem(); }

int symx_scene_source(symx_scene s, int32_t i, char* buf, int64_t cap, int64_t* len)
{
    const auto& inf = s->newton->device().info();
    if (i < 0 || i >= (int32_t)inf.size()) return 1;
    const std::string& src = inf[i].source;
    if (len) *len = (int64_t)src.size();
    if (buf && cap > 0) {
        const int64_t n = std::min<int64_t>(cap - 1, (int64_t)src.size());
        std::memcpy(buf, src.data(), (size_t)n);
        buf[n] = '\0';
    }
    return 0;
}
int symx_scene_op_counts(symx_scene s, int32_t i, int64_t* ops3)
{
    const auto& inf = s->newton->device().info();
    if (i < 0 || i >= (int32_t)inf.size()) return 1;
    ops3[0] = inf[i].ops_P;
    ops3[1] = inf[i].ops_PG;
    ops3[2] = inf[i].ops_PGH;
    return 0;
}
double symx_scene_compile_seconds(symx_scene s) { return s->newton->device().compile_seconds; }

// ---- host-only helpers ------------------------------------------------------------
int symx_host_tet4_element(int32_t material, const double* x12, const double* X12, double a, double b, double* E, double* g12, double* H144)
{
    return guard([&]() {
        std::vector<std::array<double, 3>> x(4), X(4);
        for (int i = 0; i < 4; i++)
            for (int d = 0; d < 3; d++) {
                x[i][d] = x12[3 * i + d];
                X[i][d] = X12[3 * i + d];
            }
        std::vector<int32_t> conn = { 0, 1, 2, 3 };
        double av = a, bv = b;
        auto [mws, elem] = MappedWorkspace<double>::create(conn, 4);
        std::vector<Vector> xs = mws->make_vectors(x, elem);
        std::vector<Vector> Xs = mws->make_vectors(X, elem);
        const Scalar sa = mws->make_scalar(av);
        const Scalar sb = mws->make_scalar(bv);
        const Scalar P = tet4_elastic(*mws, xs, Xs, sa, sb, material);
        std::vector<Scalar> dofs = collect_scalars(xs);
        DiffCache dc;
        const Vector g = gradient(P, dofs, dc);
        const Matrix H = gradient(g, dofs, true, dc);
        Graph* gr = mws->workspace().graph();
        // bind symbol values: maps in creation order, then the single quadrature point
        for (const DataMap& m : mws->get_maps()) {
            const double* d = m.data();
            const int32_t idx = m.connectivity_index >= 0 ? conn[m.connectivity_index] : 0;
            for (int k = 0; k < m.stride; k++) gr->set_symbol_value(m.first_symbol_idx + k, d[(size_t)idx * m.stride + k]);
        }
        const SummationMap& sm = mws->get_summation();
        for (int k = 0; k < sm.stride; k++) gr->set_symbol_value(sm.first_symbol_idx + k, sm.data[k]);
        std::vector<int32_t> ids = { P.id };
        for (const Scalar& s : g.values()) ids.push_back(s.id);
        for (const Scalar& s : H.values()) ids.push_back(s.id);
        const std::vector<double> v = gr->eval_many(ids);
        *E = v[0];
        for (int i = 0; i < 12; i++) g12[i] = v[1 + i];
        for (int i = 0; i < 144; i++) H144[i] = v[13 + i];
    });
}

int symx_host_point_triangle(const double* in, double* out)
{
    return guard([&]() {
        Workspace ws;
        Vector a = ws.make_vector(3), b = ws.make_vector(3), c = ws.make_vector(3), p = ws.make_vector(3);
        Scalar sign = ws.make_scalar();
        const Vector ap = p - a;
        const Vector ca = a - c;
        const Vector cb = b - c;
        const Vector n = ca.cross3(cb).normalized();
        const Scalar C = sign * ap.dot(n);
        std::vector<Scalar> vars = { a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], p[0], p[1], p[2] };
        const Vector g = gradient(C, vars);
        const Matrix H = hessian(C, vars);
        for (int i = 0; i < 12; i++) vars[i].set_value(in[i]);
        sign.set_value(in[12]);
        std::vector<int32_t> ids = { C.id };
        for (const Scalar& s : g.values()) ids.push_back(s.id);
        for (const Scalar& s : H.values()) ids.push_back(s.id);
        const std::vector<double> v = ws.graph()->eval_many(ids);
        for (int i = 0; i < 157; i++) out[i] = v[i];
    });
}

int symx_host_jit_tet4(const char* cache_dir, int32_t material, int64_t* code_size, int64_t* n_ops_pgh)
{
    return guard([&]() {
        std::vector<std::array<double, 3>> x(4, { 0.0, 0.0, 0.0 }), X(4, { 0.0, 0.0, 0.0 });
        std::vector<int32_t> conn = { 0, 1, 2, 3 };
        double av = 1.0, bv = 1.0;
        auto [mws, elem] = MappedWorkspace<double>::create(conn, 4);
        std::vector<Vector> xs = mws->make_vectors(x, elem);
        std::vector<Vector> Xs = mws->make_vectors(X, elem);
        const Scalar P = tet4_elastic(*mws, xs, Xs, mws->make_scalar(av), mws->make_scalar(bv), material);
        std::vector<Scalar> dofs = collect_scalars(xs);
        DiffCache dc;
        const Vector g = gradient(P, dofs, dc);
        const Matrix H = gradient(g, dofs, true, dc);
        std::vector<KernelSpec> specs = {
            KernelSpec{ "tet4_P", KernelKind::P, { P }, 12 },
            KernelSpec{ "tet4_PG", KernelKind::PG, collect_scalars({ { P }, g.values() }), 12 },
            KernelSpec{ "tet4_PGH", KernelKind::PGH, collect_scalars({ { P }, g.values(), H.values() }), 12 },
        };
        const std::string src = emit_hip_module("tet4", *mws->workspace().graph(), mws->get_maps(), mws->get_summation(), specs,
                                                symx_hip_eval_prelude());
        char key[80];
        int wc = 0;
        size_t cs = 0;
        if (symx_hip_jit_compile(src.c_str(), "gfx950", cache_dir, key, &wc, &cs) != SYMX_HIP_OK)
            throw std::runtime_error(symx_hip_last_error(nullptr));
        *code_size = (int64_t)cs;
        *n_ops_pgh = count_operations(*mws->workspace().graph(), specs[2].outputs);
    });
}

int symx_host_sphere_mesh(const symx_sphere_params* p, int32_t* nv, int32_t* ne, int32_t* nc, double* X, int32_t* tets, int32_t* contact)
{
    return guard([&]() {
        std::vector<std::array<double, 3>> Xv;
        std::vector<int32_t> t, c;
        sphere_mesh(p, Xv, t, c);
        *nv = (int32_t)Xv.size();
        *ne = (int32_t)(t.size() / 4);
        *nc = (int32_t)c.size();
        if (X) std::memcpy(X, Xv.data(), Xv.size() * 3 * sizeof(double));
        if (tets) std::memcpy(tets, t.data(), t.size() * sizeof(int32_t));
        if (contact) std::memcpy(contact, c.data(), c.size() * sizeof(int32_t));
    });
}

int symx_host_hex_element(const symx_hex_block_params* p, int32_t structured, int32_t e, double* E, double* g, double* H, int32_t* was_structured)
{
    return guard([&]() {
        const bool prev = get_structured_fem();
        set_structured_fem(structured != 0);
        symx_newton_params np;
        symx_newton_params_default(&np);
        symx_scene_s s;
        try {
            build_hex(&s, p, &np);
        }
        catch (...) {
            set_structured_fem(prev);
            throw;
        }
        set_structured_fem(prev);
        const Potential& P = *s.G->get_potentials()[0];
        const PotentialModule pm = generate_potential_module(*s.G, P);
        MappedWorkspace<double>& mws = *P.mws;
        Graph* gr = mws.workspace().graph();
        const SummationMap& sm = mws.get_summation();
        const int32_t* conn = mws.conn.data() + (size_t)e * mws.conn.stride;
        for (const DataMap& m : mws.get_maps()) {
            const int64_t item = m.connectivity_index >= 0 ? conn[m.connectivity_index] : 0;
            for (int k = 0; k < m.stride; k++) gr->set_symbol_value(m.first_symbol_idx + k, m.data()[item * m.stride + k]);
        }
        const int n = pm.info.n_dofs;
        std::fill(g, g + n, 0.0);
        std::fill(H, H + (size_t)n * n, 0.0);
        *E = 0.0;
        *was_structured = pm.structured ? 1 : 0;
        for (int32_t q = 0; q < sm.n_iterations; q++) {
            for (int k = 0; k < sm.stride; k++) gr->set_symbol_value(sm.first_symbol_idx + k, sm.data[(size_t)q * sm.stride + k]);
            if (pm.structured) {
                // the structured kernel's arithmetic, on the host
                const FemStructure& f = pm.fem;
                const int NN = f.n_nodes;
                for (const auto& rec : f.records) {
                    const DataMap& base = mws.get_maps()[rec.map_base];
                    double J[9] = { 0 };
                    for (int a = 0; a < NN; a++) {
                        const double* xa = base.data() + (size_t)conn[rec.cols[a]] * 3;
                        const double* d = &f.dN[((size_t)q * NN + a) * 3];
                        for (int i = 0; i < 3; i++)
                            for (int k = 0; k < 3; k++) J[i * 3 + k] += xa[i] * d[k];
                    }
                    for (int k = 0; k < 9; k++) gr->set_symbol_value(rec.J[k], J[k]);
                }
                std::vector<int32_t> ids = { f.f[0].id };
                for (const Scalar& x : f.P) ids.push_back(x.id);
                for (const Scalar& x : f.A) ids.push_back(x.id);
                const std::vector<double> v = gr->eval_many(ids);
                auto A = [&](int r, int c) {
                    const int lo = std::min(r, c), hi = std::max(r, c);
                    return v[10 + lo * 9 - (lo * (lo - 1)) / 2 + (hi - lo)];
                };
                *E += v[0];
                for (int a = 0; a < NN; a++) {
                    const double* da = &f.dN[((size_t)q * NN + a) * 3];
                    const int A_ = f.block_of_node[a];
                    for (int i = 0; i < 3; i++) g[3 * A_ + i] += v[1 + i * 3] * da[0] + v[2 + i * 3] * da[1] + v[3 + i * 3] * da[2];
                    for (int b = 0; b < NN; b++) {
                        const double* db = &f.dN[((size_t)q * NN + b) * 3];
                        const int B_ = f.block_of_node[b];
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) {
                                double acc = 0.0;
                                for (int k = 0; k < 3; k++)
                                    for (int l = 0; l < 3; l++) acc += da[k] * A(i * 3 + k, j * 3 + l) * db[l];
                                H[(size_t)(3 * A_ + i) * n + 3 * B_ + j] += acc;
                            }
                    }
                }
            }
            else {
                std::vector<int32_t> ids;
                for (const Scalar& x : pm.specs[2].outputs) ids.push_back(x.id);
                const std::vector<double> v = gr->eval_many(ids);
                *E += v[0];
                for (int i = 0; i < n; i++) g[i] += v[1 + i];
                for (size_t i = 0; i < (size_t)n * n; i++) H[i] += v[1 + n + i];
            }
        }
    });
}

int symx_host_hex_jit(const symx_hex_block_params* p, const char* cache_dir, int32_t* n_potentials, int64_t* code_bytes, int32_t* n_structured)
{
    return guard([&]() {
        symx_newton_params np;
        symx_newton_params_default(&np);
        symx_scene_s s;
        build_hex(&s, p, &np);
        *n_potentials = 0;
        *code_bytes = 0;
        *n_structured = 0;
        for (const auto& pp : s.G->get_potentials()) {
            const PotentialModule pm = generate_potential_module(*s.G, *pp);
            char key[80];
            int wc = 0;
            size_t cs = 0;
            if (symx_hip_jit_compile(pm.source.c_str(), "gfx950", cache_dir, key, &wc, &cs) != SYMX_HIP_OK)
                throw std::runtime_error(std::string("JIT of ") + pp->name + ": " + symx_hip_last_error(nullptr));
            *n_potentials += 1;
            *code_bytes += (int64_t)cs;
            *n_structured += pm.structured ? 1 : 0;
        }
    });
}

int symx_host_scene_jit(const symx_tet4_block_params* p, const char* cache_dir, int32_t* n_potentials, int64_t* code_bytes, int64_t* ops_adjoint)
{
    return guard([&]() {
        symx_newton_params np;
        symx_newton_params_default(&np);
        symx_scene_s s;
        build_scene(&s, p, &np);
        *n_potentials = 0;
        *code_bytes = 0;
        *ops_adjoint = 0;
        for (const auto& pp : s.G->get_potentials()) {
            const PotentialModule pm = generate_potential_module(*s.G, *pp);
            char key[80];
            int wc = 0;
            size_t cs = 0;
            if (symx_hip_jit_compile(pm.source.c_str(), "gfx950", cache_dir, key, &wc, &cs) != SYMX_HIP_OK)
                throw std::runtime_error(std::string("JIT of ") + pp->name + ": " + symx_hip_last_error(nullptr));
            *n_potentials += 1;
            *code_bytes += (int64_t)cs;
            *ops_adjoint += pm.info.ops_adjoint;
        }
    });
}

int symx_host_scene_element(const symx_tet4_block_params* p, int32_t chain_rule, int32_t potential, int32_t e, double* E, double* g,
                            double* H, int32_t* n_dofs, int32_t* was_factorised)
{
    return guard([&]() {
        symx_newton_params np;
        symx_newton_params_default(&np);
        symx_scene_s s;
        build_scene(&s, p, &np);
        const auto& pots = s.G->get_potentials();
        if (potential < 0 || potential >= (int32_t)pots.size()) throw std::runtime_error("potential index out of range");
        const Potential& P = *pots[potential];
        const bool prev = get_chain_rule();
        set_chain_rule(chain_rule != 0);
        PotentialModule pm;
        try {
            pm = generate_potential_module(*s.G, P);
        }
        catch (...) {
            set_chain_rule(prev);
            throw;
        }
        set_chain_rule(prev);
        const int n = pm.info.n_dofs;
        if (n > 36) throw std::runtime_error("element has more than 36 DoFs");
        MappedWorkspace<double>& mws = *P.mws;
        if (e < 0 || e >= mws.conn.n_elements()) throw std::runtime_error("element index out of range");
        Graph* gr = mws.workspace().graph();
        const SummationMap& sm = mws.get_summation();
        if (sm.n_iterations > 1) throw std::runtime_error("summed (multi-point) potential");
        const int32_t* conn = mws.conn.data() + (size_t)e * mws.conn.stride;
        for (const DataMap& m : mws.get_maps()) {
            const int64_t item = m.connectivity_index >= 0 ? conn[m.connectivity_index] : 0;
            for (int k = 0; k < m.stride; k++) gr->set_symbol_value(m.first_symbol_idx + k, m.data()[item * m.stride + k]);
        }
        for (int k = 0; k < sm.stride && sm.n_iterations == 1; k++) gr->set_symbol_value(sm.first_symbol_idx + k, sm.data[k]);
        std::vector<int32_t> ids;
        for (const Scalar& x : pm.specs[2].outputs) ids.push_back(x.id);
        const std::vector<double> v = gr->eval_many(ids);
        *E = v[0];
        std::copy(v.begin() + 1, v.begin() + 1 + n, g);
        std::copy(v.begin() + 1 + n, v.begin() + 1 + n + (size_t)n * n, H);
        *n_dofs = n;
        *was_factorised = pm.factorised ? 1 : 0;
    });
}

int symx_host_scene_adjoint_element(const symx_tet4_block_params* p, const double* log_mu, int32_t e, const double* lambda, double* out)
{
    return guard([&]() {
        if (!p->log_mu_field) throw std::runtime_error("scene has no log(mu) parameters");
        symx_newton_params np;
        symx_newton_params_default(&np);
        symx_scene_s s;
        build_scene(&s, p, &np);
        std::memcpy(s.log_mu.data(), log_mu, s.log_mu.size() * sizeof(double));
        const Potential& P = *s.G->get_potentials()[0];
        const PotentialModule pm = generate_potential_module(*s.G, P);
        const KernelSpec& ks = pm.specs.back();
        if (ks.kind != KernelKind::PGP) throw std::runtime_error("elastic potential has no adjoint kernel");
        MappedWorkspace<double>& mws = *P.mws;
        Graph* g = mws.workspace().graph();
        const int32_t* conn = mws.conn.data() + (size_t)e * mws.conn.stride;
        // Gather like the generated kernel: data maps through their connectivity column,
        // adjoint symbols where their DoF map reads (the DoF set starts at offset 0 here).
        for (const DataMap& m : mws.get_maps()) {
            const int64_t item = m.connectivity_index >= 0 ? conn[m.connectivity_index] : 0;
            for (int k = 0; k < m.stride; k++) g->set_symbol_value(m.first_symbol_idx + k, m.data()[item * m.stride + k]);
        }
        for (const auto& a : ks.adjoint_symbols) {
            const DataMap& m = mws.get_maps()[a[1]];
            g->set_symbol_value(a[0], lambda[(int64_t)conn[m.connectivity_index] * m.stride + a[2]]);
        }
        std::vector<int32_t> ids;
        for (const Scalar& v : ks.outputs) ids.push_back(v.id);
        // summation (quadrature) symbols: accumulate over the iterations like the kernel
        const SummationMap& sm = mws.get_summation();
        const int32_t nq = std::max<int32_t>(sm.n_iterations, 1);
        for (size_t i = 0; i < ids.size(); i++) out[i] = 0.0;
        for (int32_t q = 0; q < nq; q++) {
            for (int k = 0; k < sm.stride; k++) g->set_symbol_value(sm.first_symbol_idx + k, sm.data[(size_t)q * sm.stride + k]);
            const std::vector<double> v = g->eval_many(ids);
            for (size_t i = 0; i < v.size(); i++) out[i] += v[i];
        }
    });
}

} // extern "C"
