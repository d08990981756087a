/*
 * CPU oracle of the hot path - TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline; the product
 * (symx_amd/) never calls it. It restates the reference algorithm in plain C
 * (FP64, OpenMP over elements and rows), each function citing the SymX code it
 * follows (paths relative to the SymX repository root):
 *
 *   oracle_nh_tet4        examples/elasticity_potentials.h:7-13 with the Tet4
 *                         fem_integrator of symx/src/solver/fem_integrators.cpp:52-58,121-133
 *                         (closed-form first/second Piola derivatives, derived by hand -
 *                         independent of the symbolic engine under test)
 *   oracle_project        symx/src/solver/second_order/project_to_PD.cpp:13-32
 *                         (eigen-decomposition by cyclic Jacobi; clamp / mirror; rebuild
 *                         only when some eigenvalue is below eps)
 *   assembly              symx/src/solver/second_order/ElementHessians.cpp:224-256
 *   block Jacobi inverse  symx/extern/BlockedSparseMatrix/.../BlockedSparseMatrix.h:1147-1214
 *   oracle_pcg_bsr        symx/extern/BlockedSparseMatrix/.../solve_pcg.h:83-232
 *   oracle_newton_solve   symx/src/solver/NewtonsMethod.cpp:28-252 (ProjectedNewton / Newton),
 *                         :388-457 (BDPCG + forcing), :459-641 (line search)
 *   oracle_adjoint_logmu  no SymX counterpart: unprojected assembly + PCG above, and the
 *                         closed-form mixed partial of an energy linear in mu_e
 *
 * Parity pins (tests/): oracle_nh_tet4 against central differences of an
 * independent numpy energy and against the reference's own Compiled<double>
 * (oracle/_ref); oracle_project against numpy.linalg.eigh; PCG against
 * scipy.sparse.linalg.spsolve and textbook PCG; Newton against the reference
 * NewtonsMethod run through oracle/_ref on the same inputs; the adjoint against
 * central differences of the loss through re-solved equilibria, a dense numpy solve,
 * and the reference's own pieces composed in oracle/_ref (ref_driver adjoint).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_API __attribute__((visibility("default")))

/* ---- small dense helpers ------------------------------------------------- */
static double det3(const double* m) /* row-major 3x3 */
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
static void inv3(const double* m, double* r)
{
    const double d = det3(m);
    const double id = 1.0 / d;
    r[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    r[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    r[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    r[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    r[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    r[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

/* ---- Neo-Hookean Tet4 element (closed form) ------------------------------ */
/* E = vol * [ mu/2 (|F|^2 - 3) - mu ln J + lambda/2 ln^2 J ],  F = Ds Dm^-1, vol = det(Dm)/6.
 * P = mu F + (lambda ln J - mu) F^-T,
 * dP[dF] = mu dF + (mu - lambda ln J) F^-T dF^T F^-T + lambda (F^-T : dF) F^-T. */
ORACLE_API int oracle_nh_tet4(const double* x, const double* X, double lambda, double mu, double* E, double* g, double* H)
{
    double Ds[9], Dm[9], Dminv[9], F[9], Finv[9], FinvT[9];
    for (int i = 0; i < 3; i++) {
        for (int k = 0; k < 3; k++) {
            Ds[3 * i + k] = x[3 * (k + 1) + i] - x[i];
            Dm[3 * i + k] = X[3 * (k + 1) + i] - X[i];
        }
    }
    inv3(Dm, Dminv);
    const double vol = det3(Dm) / 6.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += Ds[3 * i + k] * Dminv[3 * k + j];
            F[3 * i + j] = s;
        }
    const double J = det3(F);
    if (!(J > 0.0)) {
        if (E) *E = INFINITY;
        return 1;
    }
    inv3(F, Finv);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) FinvT[3 * i + j] = Finv[3 * j + i];
    const double lnJ = log(J);
    double IC = 0.0;
    for (int i = 0; i < 9; i++) IC += F[i] * F[i];
    if (E) *E = vol * (0.5 * mu * (IC - 3.0) - mu * lnJ + 0.5 * lambda * lnJ * lnJ);
    /* dF_ij / dx_(a,d): a >= 1: delta_id Dminv[a-1][j]; a = 0: -delta_id sum_k Dminv[k][j] */
    double B[9][12];
    memset(B, 0, sizeof(B));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int a = 1; a < 4; a++) {
                B[3 * i + j][3 * a + i] = Dminv[3 * (a - 1) + j];
                s += Dminv[3 * (a - 1) + j];
            }
            B[3 * i + j][0 + i] = -s;
        }
    if (g) {
        double P[9];
        for (int i = 0; i < 9; i++) P[i] = mu * F[i] + (lambda * lnJ - mu) * FinvT[i];
        for (int c = 0; c < 12; c++) {
            double s = 0.0;
            for (int r = 0; r < 9; r++) s += B[r][c] * P[r];
            g[c] = vol * s;
        }
    }
    if (H) {
        double K[9][9]; /* K[r][c] = d P_r / d F_c */
        for (int c = 0; c < 9; c++) {
            double dF[9] = { 0 };
            dF[c] = 1.0;
            double trm = 0.0; /* F^-T : dF */
            for (int t = 0; t < 9; t++) trm += FinvT[t] * dF[t];
            /* M = F^-T dF^T F^-T */
            double dFT[9], T1[9], M[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) dFT[3 * i + j] = dF[3 * j + i];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++) s += FinvT[3 * i + k] * dFT[3 * k + j];
                    T1[3 * i + j] = s;
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++) s += T1[3 * i + k] * FinvT[3 * k + j];
                    M[3 * i + j] = s;
                }
            for (int r = 0; r < 9; r++) K[r][c] = mu * dF[r] + (mu - lambda * lnJ) * M[r] + lambda * trm * FinvT[r];
        }
        for (int a = 0; a < 12; a++)
            for (int b = 0; b < 12; b++) {
                double s = 0.0;
                for (int r = 0; r < 9; r++) {
                    if (B[r][a] == 0.0) continue;
                    double t = 0.0;
                    for (int c = 0; c < 9; c++) t += K[r][c] * B[c][b];
                    s += B[r][a] * t;
                }
                H[12 * a + b] = vol * s;
            }
    }
    return 0;
}

/* ---- Neo-Hookean Hex8 / Hex27 / Tet10 elements (closed form, isoparametric) */
/* Node order, shape functions and quadrature of fem_integrators.h:56-150 and
 * fem_integrators.cpp:60-116. Hexahedra: per axis the node sits at -1, 0 or +1 (0, 1, 2
 * below); Hex8: 0.5 (1 -+ t); Hex27: t(t-1), 1 - t^2, t(t+1) scaled by 1/8, 1/4, 1/2, 1 for
 * nodes with 0, 1, 2, 3 mid-axis positions. Rules: 2x2x2 Gauss (weight 1) and the 27-point
 * rule with the reference's 10-digit constants, in the reference's order. Tet10 (quadratic
 * tetrahedron, barycentric L0 = 1 - u - v - w, L1 = u, L2 = v, L3 = w): vertices
 * Li (2 Li - 1), then mid-edges 4 Li Lj in the order 01 12 02 03 13 23; the reference's
 * 4-point rule with its 8-digit constants (weights 0.041666667, summing to 1/6 + 2e-9). */
static const int8_t hex_nodes[27][3] = {
    { 0, 0, 0 }, { 2, 0, 0 }, { 2, 2, 0 }, { 0, 2, 0 }, { 0, 0, 2 }, { 2, 0, 2 }, { 2, 2, 2 }, { 0, 2, 2 },
    { 1, 0, 0 }, { 2, 1, 0 }, { 1, 2, 0 }, { 0, 1, 0 }, { 1, 0, 2 }, { 2, 1, 2 }, { 1, 2, 2 }, { 0, 1, 2 },
    { 0, 0, 1 }, { 2, 0, 1 }, { 2, 2, 1 }, { 0, 2, 1 },
    { 0, 1, 1 }, { 2, 1, 1 }, { 1, 2, 1 }, { 1, 0, 1 }, { 1, 1, 0 }, { 1, 1, 2 },
    { 1, 1, 1 } };

static int iso_rule(int nn, double (*q)[4])
{
    if (nn == 10) {
        const double a = 0.58541020, b = 0.13819660, w = 0.041666667;
        const double t[4][4] = { { w, a, b, b }, { w, b, a, b }, { w, b, b, a }, { w, b, b, b } };
        memcpy(q, t, sizeof(t));
        return 4;
    }
    if (nn == 8) {
        const double p = 1.0 / sqrt(3.0);
        const double pts[8][3] = { { -p, -p, -p }, { p, -p, -p }, { p, p, -p }, { -p, p, -p }, { -p, -p, p }, { p, -p, p }, { p, p, p }, { -p, p, p } };
        for (int i = 0; i < 8; i++) {
            q[i][0] = 1.0;
            for (int d = 0; d < 3; d++) q[i][1 + d] = pts[i][d];
        }
        return 8;
    }
    const double p = 0.7745966692, wc = 0.1714677641, we = 0.2743484225, wf = 0.4389574760, wm = 0.7023319616;
    const double t[27][4] = { { wc, -p, -p, -p }, { wc, p, -p, -p }, { wc, -p, p, -p }, { wc, p, p, -p }, { wc, -p, -p, p }, { wc, p, -p, p },
                              { wc, -p, p, p }, { wc, p, p, p }, { we, 0.0, -p, -p }, { we, -p, 0.0, -p }, { we, p, 0.0, -p }, { we, 0.0, p, -p },
                              { we, -p, -p, 0.0 }, { we, p, -p, 0.0 }, { we, -p, p, 0.0 }, { we, p, p, 0.0 }, { we, 0.0, -p, p }, { we, -p, 0.0, p },
                              { we, p, 0.0, p }, { we, 0.0, p, p }, { wf, 0.0, 0.0, -p }, { wf, 0.0, -p, 0.0 }, { wf, -p, 0.0, 0.0 },
                              { wf, p, 0.0, 0.0 }, { wf, 0.0, p, 0.0 }, { wf, 0.0, 0.0, p }, { wm, 0.0, 0.0, 0.0 } };
    memcpy(q, t, sizeof(t));
    return 27;
}

/* dN_a / dxi at xi (dN [nn][3]) */
static void iso_dshape(int nn, const double* xi, double* dN)
{
    if (nn == 10) {
        const double L[4] = { 1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2] };
        const double dL[4][3] = { { -1.0, -1.0, -1.0 }, { 1.0, 0.0, 0.0 }, { 0.0, 1.0, 0.0 }, { 0.0, 0.0, 1.0 } };
        const int edge[6][2] = { { 0, 1 }, { 1, 2 }, { 0, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) dN[3 * i + k] = (4.0 * L[i] - 1.0) * dL[i][k];
        for (int m = 0; m < 6; m++) {
            const int i = edge[m][0], j = edge[m][1];
            for (int k = 0; k < 3; k++) dN[3 * (4 + m) + k] = 4.0 * (dL[i][k] * L[j] + L[i] * dL[j][k]);
        }
        return;
    }
    for (int a = 0; a < nn; a++) {
        double f[3], df[3];
        int n_mid = 0;
        for (int d = 0; d < 3; d++) {
            const int pos = hex_nodes[a][d];
            const double t = xi[d];
            if (nn == 8) {
                const double s = pos ? 1.0 : -1.0;
                f[d] = 0.5 * (1.0 + s * t);
                df[d] = 0.5 * s;
            }
            else if (pos == 0) {
                f[d] = t * (t - 1.0);
                df[d] = 2.0 * t - 1.0;
            }
            else if (pos == 2) {
                f[d] = t * (t + 1.0);
                df[d] = 2.0 * t + 1.0;
            }
            else {
                f[d] = 1.0 - t * t;
                df[d] = -2.0 * t;
                n_mid++;
            }
        }
        const double scale = nn == 8 ? 1.0 : (n_mid == 0 ? 0.125 : (n_mid == 1 ? 0.25 : (n_mid == 2 ? 0.5 : 1.0)));
        dN[3 * a + 0] = scale * df[0] * f[1] * f[2];
        dN[3 * a + 1] = scale * f[0] * df[1] * f[2];
        dN[3 * a + 2] = scale * f[0] * f[1] * df[2];
    }
}

/* E = sum_q w_q det(JX_q) psi(F_q), F = Jx JX^-1, psi the Neo-Hookean density above.
 * dF_ij/dx_(a,i) = G_a[j] = sum_k dN_a[k] JX^-1[k][j]. x, X: [nn][3]; g: [3nn]; H: [3nn][3nn]. */
ORACLE_API int oracle_nh_hex(int nn, const double* x, const double* X, double lambda, double mu, double* E, double* g, double* H)
{
    if (nn != 8 && nn != 10 && nn != 27) return 2;
    const int n = 3 * nn;
    double q[27][4];
    const int nq = iso_rule(nn, q);
    double dN[27 * 3], G[27 * 3];
    if (E) *E = 0.0;
    if (g) memset(g, 0, sizeof(double) * n);
    if (H) memset(H, 0, sizeof(double) * n * n);
    for (int iq = 0; iq < nq; iq++) {
        iso_dshape(nn, &q[iq][1], dN);
        double Jx[9] = { 0 }, JX[9] = { 0 }, JXinv[9], F[9], Finv[9], FinvT[9];
        for (int a = 0; a < nn; a++)
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) {
                    Jx[3 * i + k] += x[3 * a + i] * dN[3 * a + k];
                    JX[3 * i + k] += X[3 * a + i] * dN[3 * a + k];
                }
        inv3(JX, JXinv);
        const double w = q[iq][0] * det3(JX);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += Jx[3 * i + k] * JXinv[3 * k + j];
                F[3 * i + j] = s;
            }
        for (int a = 0; a < nn; a++)
            for (int j = 0; j < 3; j++) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += dN[3 * a + k] * JXinv[3 * k + j];
                G[3 * a + j] = s;
            }
        const double J = det3(F);
        if (!(J > 0.0)) {
            if (E) *E = INFINITY;
            return 1;
        }
        inv3(F, Finv);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) FinvT[3 * i + j] = Finv[3 * j + i];
        const double lnJ = log(J);
        double IC = 0.0;
        for (int i = 0; i < 9; i++) IC += F[i] * F[i];
        if (E) *E += w * (0.5 * mu * (IC - 3.0) - mu * lnJ + 0.5 * lambda * lnJ * lnJ);
        if (g) {
            double P[9];
            for (int i = 0; i < 9; i++) P[i] = mu * F[i] + (lambda * lnJ - mu) * FinvT[i];
            for (int a = 0; a < nn; a++)
                for (int i = 0; i < 3; i++) g[3 * a + i] += w * (P[3 * i + 0] * G[3 * a + 0] + P[3 * i + 1] * G[3 * a + 1] + P[3 * i + 2] * G[3 * a + 2]);
        }
        if (H) {
            /* K[(i,j)][(k,l)] = dP_ij / dF_kl = mu d_ik d_jl + (mu - lambda lnJ) Finv_jk Finv_li + lambda Finv_ji Finv_lk */
            double K[9][9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++)
                            K[3 * i + j][3 * k + l] = (i == k && j == l ? mu : 0.0) + (mu - lambda * lnJ) * Finv[3 * j + k] * Finv[3 * l + i] +
                                                      lambda * Finv[3 * j + i] * Finv[3 * l + k];
            for (int a = 0; a < nn; a++)
                for (int b = 0; b < nn; b++)
                    for (int i = 0; i < 3; i++)
                        for (int k = 0; k < 3; k++) {
                            double s = 0.0;
                            for (int j = 0; j < 3; j++)
                                for (int l = 0; l < 3; l++) s += G[3 * a + j] * K[3 * i + j][3 * k + l] * G[3 * b + l];
                            H[(3 * a + i) * n + 3 * b + k] += w * s;
                        }
        }
    }
    return 0;
}

/* element energy of a scene element (Tet4, Tet10 or hexahedron) */
static int scene_element(int npe, const double* xe, const double* Xe, double lam, double mu, double* E, double* g, double* H)
{
    return npe == 4 ? oracle_nh_tet4(xe, Xe, lam, mu, E, g, H) : oracle_nh_hex(npe, xe, Xe, lam, mu, E, g, H);
}

/* ---- PD projection (project_to_PD.cpp:13-32) ---------------------------- */
/* Eigen-decomposition of a symmetric n x n matrix by cyclic Jacobi; eigenvalues
 * below eps become eps (clamp) or -lambda (mirror); the matrix is rebuilt as
 * V f(Lambda) V^T only when some eigenvalue was below eps. Returns 1 if rebuilt. */
ORACLE_API int oracle_project(double* A, int n, double eps, int mirror)
{
    double* a = (double*)malloc(sizeof(double) * n * n);
    double* v = (double*)malloc(sizeof(double) * n * n);
    double* lam = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n * n; i++) a[i] = 0.5 * (A[i] + A[(i % n) * n + i / n]);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[i * n + j] = (i == j) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = 0; i < n * n; i++) fro += a[i] * a[i];
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        if (off <= 1e-32 * fro) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) { /* A <- A J (columns p, q) */
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) { /* A <- J^T A (rows p, q) */
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    int changed = 0;
    for (int i = 0; i < n; i++) {
        lam[i] = a[i * n + i];
        if (lam[i] < eps) {
            changed = 1;
            lam[i] = mirror ? -lam[i] : eps;
        }
    }
    if (changed) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double s = 0.0;
                for (int k = 0; k < n; k++) s += v[i * n + k] * lam[k] * v[j * n + k];
                A[i * n + j] = s;
            }
    }
    free(a);
    free(v);
    free(lam);
    return changed;
}

ORACLE_API int oracle_project_many(double* A, int64_t count, int n, double eps, int mirror, uint8_t* changed)
{
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int64_t i = 0; i < count; i++) {
        const int c = oracle_project(A + i * n * n, n, eps, mirror);
        if (changed) changed[i] = (uint8_t)c;
        total += c;
    }
    return (int)total;
}

/* ---- block CSR, block-Jacobi PCG (solve_pcg.h:83-232) -------------------- */
typedef struct oracle_pcg_info {
    int32_t n_iterations;
    int32_t converged;
    int32_t found_indefiniteness;
    int32_t pad;
    double error;
} oracle_pcg_info;

static void bsr_spmv(int nb, const int64_t* rp, const int32_t* cols, const double* vals, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nb; r++) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int64_t k = rp[r]; k < rp[r + 1]; k++) {
            const double* b = vals + 9 * k;
            const double* xv = x + 3 * (int64_t)cols[k];
            a0 += b[0] * xv[0] + b[1] * xv[1] + b[2] * xv[2];
            a1 += b[3] * xv[0] + b[4] * xv[1] + b[5] * xv[2];
            a2 += b[6] * xv[0] + b[7] * xv[1] + b[8] * xv[2];
        }
        y[3 * (int64_t)r] = a0;
        y[3 * (int64_t)r + 1] = a1;
        y[3 * (int64_t)r + 2] = a2;
    }
}
static double vdot(int64_t n, const double* a, const double* b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
/* closed-form inverse of the symmetric diagonal blocks (BlockedSparseMatrix.h:1198-1214) */
static void block_jacobi(int nb, const int64_t* rp, const int32_t* cols, const double* vals, double* dinv)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nb; r++) {
        memset(dinv + 9 * (int64_t)r, 0, 9 * sizeof(double));
        for (int64_t k = rp[r]; k < rp[r + 1]; k++) {
            if (cols[k] == r) {
                inv3(vals + 9 * k, dinv + 9 * (int64_t)r);
                break;
            }
        }
    }
}

ORACLE_API int oracle_pcg_bsr(int nb, const int64_t* rp, const int32_t* cols, const double* vals, const double* b_in, double* x_out,
                              double abs_tol, double rel_tol, int max_iter, int stop_on_indef, oracle_pcg_info* info)
{
    const int64_t n = 3 * (int64_t)nb;
    double* dinv = (double*)malloc(sizeof(double) * 9 * nb);
    double* x = (double*)calloc(n, sizeof(double));
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* Ap = (double*)malloc(sizeof(double) * n);
    block_jacobi(nb, rp, cols, vals, dinv);
    memset(info, 0, sizeof(*info));
    const double bb = vdot(n, b_in, b_in);
    if (bb < abs_tol * abs_tol) {
        memset(x_out, 0, sizeof(double) * n);
        info->converged = 1;
        goto done;
    }
    bsr_spmv(nb, rp, cols, vals, x, Ap);
    for (int64_t i = 0; i < n; i++) r[i] = b_in[i] - Ap[i];
    double rr = vdot(n, r, r);
    double err = sqrt(rr / bb);
    const double err0 = err;
    if (err < abs_tol) {
        memcpy(x_out, x, sizeof(double) * n);
        info->converged = 1;
        info->error = err;
        goto done;
    }
#define APPLY_PRECOND()                                                                                        \
    for (int row = 0; row < nb; row++) {                                                                       \
        const double* m = dinv + 9 * (int64_t)row;                                                              \
        const double* rv = r + 3 * (int64_t)row;                                                                \
        for (int i = 0; i < 3; i++) z[3 * (int64_t)row + i] = m[3 * i] * rv[0] + m[3 * i + 1] * rv[1] + m[3 * i + 2] * rv[2]; \
    }
    APPLY_PRECOND();
    memcpy(p, z, sizeof(double) * n);
    double rz = vdot(n, r, z);
    for (int it = 1; it <= max_iter; it++) {
        bsr_spmv(nb, rp, cols, vals, p, Ap);
        const double pAp = vdot(n, p, Ap);
        if (pAp <= 0.0) {
            info->found_indefiniteness = 1;
            if (stop_on_indef) {
                memcpy(x_out, x, sizeof(double) * n);
                info->converged = 0;
                info->n_iterations = it;
                info->error = err;
                goto done;
            }
        }
        const double alpha = rz / pAp;
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        rr = vdot(n, r, r);
        err = sqrt(rr / bb);
        if (err < abs_tol || err / err0 < rel_tol) {
            memcpy(x_out, x, sizeof(double) * n);
            info->converged = 1;
            info->n_iterations = it;
            info->error = err;
            goto done;
        }
        APPLY_PRECOND();
        const double rz_old = rz;
        rz = vdot(n, r, z);
        const double beta = rz / rz_old;
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    memcpy(x_out, x, sizeof(double) * n);
    info->converged = 0;
    info->n_iterations = max_iter;
    info->error = err;
done:
    free(dinv);
    free(x);
    free(r);
    free(z);
    free(p);
    free(Ap);
    return 0;
}

/* ---- Tet4 block problem (the scene of symx_scenes.h) ---------------------- */
typedef struct oracle_scene {
    int32_t nv, ne, nbc;
    int32_t has_inertia;
    const int32_t* tets;     /* [ne][4] */
    const double* X;         /* [nv][3] rest positions */
    double* x;               /* [nv][3] DoFs (in/out) */
    const double* x0;        /* [nv][3] */
    const double* v0;        /* [nv][3] */
    const double* mass;      /* [nv] */
    const int32_t* bc_vertex;/* [nbc] */
    const double* bc_target; /* [nbc][3] */
    const double* mu_e;      /* [ne] per-element mu (NULL: mu) */
    double lambda_over_mu;   /* used with mu_e */
    double lambda, mu, dt, k;
    double gravity[3];
    /* ground contact: IPC log-barrier per contact vertex against z = 0, active below dhat,
     * and the plane step filter with factor step_eta (0: off) */
    int32_t ncontact;
    int32_t pad_;
    const int32_t* contact_vertex;
    double kappa, dhat, step_eta;
    /* element type of `tets`: 4 (Tet4, also when 0), 8 (Hex8), 10 (Tet10) or 27 (Hex27) */
    int32_t nodes_per_element;
    int32_t pad2_;
} oracle_scene;

#define NPE(S) ((S)->nodes_per_element > 0 ? (S)->nodes_per_element : 4)

/* b(d) = -kappa (d - dhat)^2 ln(d / dhat) and its first two derivatives, 0 < d < dhat */
static void barrier(double kappa, double dhat, double d, double* b, double* db, double* ddb)
{
    const double l = log(d / dhat), e = d - dhat;
    *b = -kappa * e * e * l;
    *db = -kappa * (2.0 * e * l + e * e / d);
    *ddb = -kappa * (2.0 * l + 4.0 * e / d - e * e / (d * d));
}

typedef struct oracle_newton_params {
    int32_t precision;
    int32_t projection_mode;
    int32_t use_mirroring;
    int32_t max_iterations;
    double residual_tolerance_abs;
    double residual_tolerance_rel;
    double cg_rel_tolerance;
    double cg_abs_tolerance;
    int32_t cg_max_iterations;
    double projection_eps;
    double step_tolerance;
} oracle_newton_params;

typedef struct oracle_stats {
    int32_t result;
    int32_t newton_iterations;
    int32_t cg_iterations;
    int32_t ls_bt_iterations;
    double final_energy;
    double final_residual;
    double first_residual;
    double seconds_eval;
    double seconds_linear;
} oracle_stats;

typedef struct pattern {
    int nb;
    int64_t* rp;
    int32_t* cols;
    int64_t nnzb;
    int64_t* elem_slot; /* [ne][npe * npe] block slot of (a, b) */
    int64_t* bc_slot;   /* [nbc] */
    int64_t* v_slot;    /* [nv] diagonal slot */
} pattern;

static int cmp_int(const void* a, const void* b) { return (*(const int32_t*)a > *(const int32_t*)b) - (*(const int32_t*)a < *(const int32_t*)b); }

static int64_t find_slot(const pattern* P, int r, int c)
{
    int64_t lo = P->rp[r], hi = P->rp[r + 1];
    while (lo < hi) {
        const int64_t m = (lo + hi) / 2;
        if (P->cols[m] < c) lo = m + 1;
        else hi = m;
    }
    return lo;
}

static void build_pattern(const oracle_scene* S, pattern* P)
{
    const int nv = S->nv;
    P->nb = nv;
    int64_t* cnt = (int64_t*)calloc(nv + 1, sizeof(int64_t));
    for (int v = 0; v < nv; v++) cnt[v + 1] += 1;
    const int npe = NPE(S);
    for (int e = 0; e < S->ne; e++)
        for (int a = 0; a < npe; a++) cnt[S->tets[npe * (int64_t)e + a] + 1] += npe;
    for (int v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
    int32_t* cand = (int32_t*)malloc(sizeof(int32_t) * cnt[nv]);
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * nv);
    for (int v = 0; v < nv; v++) {
        cur[v] = cnt[v];
        cand[cur[v]++] = v;
    }
    for (int e = 0; e < S->ne; e++)
        for (int a = 0; a < npe; a++) {
            const int r = S->tets[npe * (int64_t)e + a];
            for (int b = 0; b < npe; b++) cand[cur[r]++] = S->tets[npe * (int64_t)e + b];
        }
    P->rp = (int64_t*)malloc(sizeof(int64_t) * (nv + 1));
    P->rp[0] = 0;
    int64_t total = 0;
    for (int v = 0; v < nv; v++) {
        int32_t* b = cand + cnt[v];
        const int64_t n = cnt[v + 1] - cnt[v];
        qsort(b, n, sizeof(int32_t), cmp_int);
        int64_t u = 0;
        for (int64_t i = 0; i < n; i++)
            if (i == 0 || b[i] != b[i - 1]) b[u++] = b[i];
        total += u;
        cur[v] = u;
    }
    P->cols = (int32_t*)malloc(sizeof(int32_t) * total);
    for (int v = 0; v < nv; v++) {
        memcpy(P->cols + P->rp[v], cand + cnt[v], sizeof(int32_t) * cur[v]);
        P->rp[v + 1] = P->rp[v] + cur[v];
    }
    P->nnzb = total;
    P->elem_slot = (int64_t*)malloc(sizeof(int64_t) * npe * npe * (int64_t)(S->ne > 0 ? S->ne : 1));
    for (int e = 0; e < S->ne; e++)
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++)
                P->elem_slot[(int64_t)npe * npe * e + npe * a + b] = find_slot(P, S->tets[npe * (int64_t)e + a], S->tets[npe * (int64_t)e + b]);
    P->v_slot = (int64_t*)malloc(sizeof(int64_t) * nv);
    for (int v = 0; v < nv; v++) P->v_slot[v] = find_slot(P, v, v);
    P->bc_slot = (int64_t*)malloc(sizeof(int64_t) * (S->nbc > 0 ? S->nbc : 1));
    for (int i = 0; i < S->nbc; i++) P->bc_slot[i] = P->v_slot[S->bc_vertex[i]];
    free(cnt);
    free(cand);
    free(cur);
}
static void free_pattern(pattern* P)
{
    free(P->rp);
    free(P->cols);
    free(P->elem_slot);
    free(P->bc_slot);
    free(P->v_slot);
}

/* Energy (and optionally gradient, element Hessians). E accumulated per element
 * in index order within each potential. Returns +inf for inverted elements. */
static double scene_eval(const oracle_scene* S, const double* x, double* grad, double* He /* [ne][(3 npe)^2] or NULL */)
{
    const int ne = S->ne, nv = S->nv, npe = NPE(S), nd = 3 * NPE(S);
    double E_el = 0.0;
    double* ge = grad ? (double*)malloc(sizeof(double) * nd * (int64_t)(ne > 0 ? ne : 1)) : NULL;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : E_el) reduction(| : bad)
    for (int e = 0; e < ne; e++) {
        double xe[81], Xe[81], Ee = 0.0;
        for (int a = 0; a < npe; a++)
            for (int d = 0; d < 3; d++) {
                const int64_t v = S->tets[npe * (int64_t)e + a];
                xe[3 * a + d] = x[3 * v + d];
                Xe[3 * a + d] = S->X[3 * v + d];
            }
        const double mu = S->mu_e ? S->mu_e[e] : S->mu;
        const double lam = S->mu_e ? S->lambda_over_mu * S->mu_e[e] : S->lambda;
        bad |= scene_element(npe, xe, Xe, lam, mu, &Ee, ge ? ge + nd * (int64_t)e : NULL, He ? He + (int64_t)nd * nd * e : NULL);
        E_el += Ee;
    }
    if (bad) E_el = INFINITY;
    double E_in = 0.0;
    if (grad) memset(grad, 0, sizeof(double) * 3 * (int64_t)nv);
    if (S->has_inertia) {
        const double dt = S->dt;
        for (int v = 0; v < nv; v++) {
            double s = 0.0;
            for (int d = 0; d < 3; d++) {
                const double xh = S->x0[3 * (int64_t)v + d] + dt * S->v0[3 * (int64_t)v + d] + dt * dt * S->gravity[d];
                const double dv = x[3 * (int64_t)v + d] - xh;
                s += dv * dv;
                if (grad) grad[3 * (int64_t)v + d] += S->mass[v] * dv / (dt * dt);
            }
            E_in += 0.5 * S->mass[v] * s / (dt * dt);
        }
    }
    else if (S->gravity[0] != 0.0 || S->gravity[1] != 0.0 || S->gravity[2] != 0.0) {
        for (int v = 0; v < nv; v++)
            for (int d = 0; d < 3; d++) {
                E_in += -S->mass[v] * S->gravity[d] * x[3 * (int64_t)v + d];
                if (grad) grad[3 * (int64_t)v + d] += -S->mass[v] * S->gravity[d];
            }
    }
    double E_contact = 0.0;
    for (int c = 0; c < S->ncontact; c++) {
        const int64_t v = S->contact_vertex[c];
        const double d = x[3 * v + 2];
        if (!(d < S->dhat)) continue;
        if (d <= 0.0) {
            E_contact = INFINITY;
            continue;
        }
        double b, db, ddb;
        barrier(S->kappa, S->dhat, d, &b, &db, &ddb);
        E_contact += b;
        if (grad) grad[3 * v + 2] += db;
    }
    double E_bc = 0.0;
    for (int i = 0; i < S->nbc; i++) {
        const int64_t v = S->bc_vertex[i];
        double s = 0.0;
        for (int d = 0; d < 3; d++) {
            const double dv = x[3 * v + d] - S->bc_target[3 * i + d];
            s += dv * dv;
            if (grad) grad[3 * v + d] += S->k * dv;
        }
        E_bc += 0.5 * S->k * s;
    }
    if (grad) {
        for (int e = 0; e < ne; e++)
            for (int a = 0; a < npe; a++)
                for (int d = 0; d < 3; d++) grad[3 * (int64_t)S->tets[npe * (int64_t)e + a] + d] += ge[nd * (int64_t)e + 3 * a + d];
        free(ge);
    }
    return E_el + E_in + E_bc + E_contact;
}

/* Barrier Hessians diag(0, 0, b'') of the active contact vertices; projected ones (PD
 * projection of the 3x3 element Hessian) get eps (clamp) or 0 (mirror) on the two zero
 * eigenvalues. */
static void add_barrier_hessians(const oracle_scene* S, const pattern* P, const double* x, const unsigned char* bproj, double eps, int mirror,
                                 double* vals)
{
    for (int c = 0; c < S->ncontact; c++) {
        const int64_t v = S->contact_vertex[c];
        const double d = x[3 * v + 2];
        if (!(d < S->dhat) || d <= 0.0) continue;
        double b, db, ddb;
        barrier(S->kappa, S->dhat, d, &b, &db, &ddb);
        double* dst = vals + 9 * P->v_slot[v];
        double h[3] = { 0.0, 0.0, ddb };
        if (bproj && bproj[c]) {
            for (int k = 0; k < 3; k++)
                if (h[k] < eps) h[k] = mirror ? -h[k] : eps;
        }
        for (int k = 0; k < 3; k++) dst[4 * k] += h[k];
    }
}

ORACLE_API double oracle_scene_energy(const oracle_scene* S, const double* x, double* grad) { return scene_eval(S, x, grad, NULL); }

/* Element Hessians of the elastic potential at x: He [ne][(3 npe)^2]. */
ORACLE_API int oracle_scene_element_hessians(const oracle_scene* S, const double* x, double* He)
{
    double* g = (double*)malloc(sizeof(double) * 3 * (int64_t)S->nv);
    scene_eval(S, x, g, He);
    free(g);
    return 0;
}

static void assemble(const oracle_scene* S, const pattern* P, const double* He, double* vals)
{
    memset(vals, 0, sizeof(double) * 9 * P->nnzb);
    const int npe = NPE(S), nd = 3 * NPE(S);
    for (int e = 0; e < S->ne; e++) {
        const double* h = He + (int64_t)nd * nd * e;
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++) {
                double* dst = vals + 9 * P->elem_slot[(int64_t)npe * npe * e + npe * a + b];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) dst[3 * i + j] += h[(3 * a + i) * nd + 3 * b + j];
            }
    }
    if (S->has_inertia) {
        for (int v = 0; v < S->nv; v++) {
            double* dst = vals + 9 * P->v_slot[v];
            for (int d = 0; d < 3; d++) dst[4 * d] += S->mass[v] / (S->dt * S->dt);
        }
    }
    for (int i = 0; i < S->nbc; i++) {
        double* dst = vals + 9 * P->bc_slot[i];
        for (int d = 0; d < 3; d++) dst[4 * d] += S->k;
    }
}

static double now_seconds(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* Newton's method of NewtonsMethod.cpp:28-252 with ProjectedNewton (mode 1) or
 * plain Newton (mode 0), BDPCG with the forcing sequence and Armijo backtracking. */
ORACLE_API int oracle_newton_solve(oracle_scene* S, const oracle_newton_params* np, oracle_stats* st, double* energies, double* residuals,
                                   int max_record)
{
    const int64_t n = 3 * (int64_t)S->nv;
    pattern P;
    build_pattern(S, &P);
    double* grad = (double*)malloc(sizeof(double) * n);
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* du = (double*)malloc(sizeof(double) * n);
    double* x_ls = (double*)malloc(sizeof(double) * n);
    const int npe = NPE(S), nd = 3 * NPE(S);
    double* He = (double*)malloc(sizeof(double) * nd * nd * (int64_t)(S->ne > 0 ? S->ne : 1));
    double* vals = (double*)malloc(sizeof(double) * 9 * P.nnzb);
    unsigned char* projected = (unsigned char*)malloc((size_t)(S->ne > 0 ? S->ne : 1));
    unsigned char* bproj = (unsigned char*)malloc((size_t)(S->ncontact > 0 ? S->ncontact : 1));
    unsigned char* active = (unsigned char*)malloc((size_t)(S->nv > 0 ? S->nv : 1));
    int pdn_countdown = 0;
    double ppn_threshold = -1.0;
    memset(st, 0, sizeof(*st));
    double res0 = 1e308;
    int result = 1; /* Running */
    int it = -1;
    while (result == 1) {
        it++;
        if (it == np->max_iterations) {
            result = 3; /* TooManyIterations */
            break;
        }
        double t0 = now_seconds();
        const double E0 = scene_eval(S, S->x, grad, He);
        double r = 0.0;
        for (int64_t i = 0; i < n; i++) r = fmax(r, fabs(grad[i]));
        if (it < max_record) {
            if (energies) energies[it] = E0;
            if (residuals) residuals[it] = r;
        }
        st->final_energy = E0;
        st->final_residual = r;
        if (it == 0) {
            res0 = r;
            st->first_residual = r;
        }
        if (r < 1e-10) {
            result = 0;
            break;
        }
        if (r < np->residual_tolerance_abs) {
            result = 0;
            break;
        }
        if (it > 0 && r / res0 < np->residual_tolerance_rel) {
            result = 0;
            break;
        }
        /* Projection + solve, retried with more projection until the step descends
         * (NewtonsMethod.cpp:133-185, _project_and_assemble :254-352, _increase/_decrease
         * :354-386). Modes: 0 Newton, 1 ProjectedNewton, 2 ProjectOnDemand, 3 Progressive
         * (default countdown 4, tightening 0.5, release 2). */
        memset(projected, 0, (size_t)(S->ne > 0 ? S->ne : 1));
        memset(bproj, 0, (size_t)(S->ncontact > 0 ? S->ncontact : 1));
        double dug = 0.0;
        for (;;) {
            int all_projected = 0;
            const int mode = np->projection_mode;
            if (mode == 1 || (mode == 2 && pdn_countdown > 0)) {
#pragma omp parallel for schedule(dynamic, 64)
                for (int e = 0; e < S->ne; e++) {
                    if (!projected[e]) oracle_project(He + (int64_t)nd * nd * e, nd, np->projection_eps, np->use_mirroring);
                    projected[e] = 1;
                }
                memset(bproj, 1, (size_t)(S->ncontact > 0 ? S->ncontact : 1));
                all_projected = 1;
            }
            else if (mode == 3 && ppn_threshold >= 0.0) {
                if (ppn_threshold > 0.0 && ppn_threshold < 1e-12) ppn_threshold = 0.0;
                int n_active = 0;
                for (int v = 0; v < S->nv; v++) {
                    const double m = fmax(fabs(grad[3 * (int64_t)v]), fmax(fabs(grad[3 * (int64_t)v + 1]), fabs(grad[3 * (int64_t)v + 2])));
                    active[v] = m >= ppn_threshold;
                    n_active += active[v];
                }
#pragma omp parallel for schedule(dynamic, 64)
                for (int e = 0; e < S->ne; e++) {
                    if (projected[e]) continue;
                    const int32_t* t = S->tets + npe * (int64_t)e;
                    int touch = 0;
                    for (int a = 0; a < npe; a++) touch |= active[t[a]];
                    if (touch) {
                        oracle_project(He + (int64_t)nd * nd * e, nd, np->projection_eps, np->use_mirroring);
                        projected[e] = 1;
                    }
                }
                for (int c = 0; c < S->ncontact; c++)
                    if (active[S->contact_vertex[c]]) bproj[c] = 1;
                all_projected = n_active == S->nv;
            }
            assemble(S, &P, He, vals);
            add_barrier_hessians(S, &P, S->x, bproj, np->projection_eps, np->use_mirroring, vals);
            st->seconds_eval += now_seconds() - t0;
            t0 = now_seconds();
            for (int64_t i = 0; i < n; i++) rhs[i] = -grad[i];
            const double forcing = fmin(1e-2, r * fmin(0.5, sqrt(r)));
            const double abs_tol = fmax(forcing, np->cg_abs_tolerance);
            oracle_pcg_info info;
            oracle_pcg_bsr(P.nb, P.rp, P.cols, vals, rhs, du, abs_tol, np->cg_rel_tolerance, np->cg_max_iterations, 1, &info);
            st->cg_iterations += info.n_iterations;
            st->seconds_linear += now_seconds() - t0;
            t0 = now_seconds();
            const int can_project_more = mode != 0 && !all_projected;
            if (!info.converged && !can_project_more) {
                result = 5; /* LinearSystemSolveFailure */
                break;
            }
            int descends = 0;
            if (info.converged) {
                dug = vdot(n, du, grad);
                descends = dug < 0.0;
                if (!descends && !can_project_more) {
                    result = 7; /* StepDoesNotDescend */
                    break;
                }
            }
            if (info.converged && descends) break;
            if (mode == 2) pdn_countdown = 4;
            if (mode == 3) {
                if (ppn_threshold < 0.0) ppn_threshold = r;
                ppn_threshold *= 0.5;
            }
        }
        if (result != 1) break;
        if (np->projection_mode == 2) pdn_countdown--;
        if (np->projection_mode == 3) ppn_threshold *= 2.0;
        double du_max = 0.0;
        for (int64_t i = 0; i < n; i++) du_max = fmax(du_max, fabs(du[i]));
        if (du_max < np->step_tolerance) {
            result = 0;
            break;
        }
        /* plane step filter (max_allowed_step, NewtonsMethod.cpp:494-506) */
        if (S->step_eta > 0.0) {
            double t = 1.0;
            for (int c = 0; c < S->ncontact; c++) {
                const int64_t k = 3 * (int64_t)S->contact_vertex[c] + 2;
                if (du[k] < 0.0) t = fmin(t, fmax(0.0, S->step_eta * S->x[k] / -du[k]));
            }
            if (t < 1.0) {
                for (int64_t i = 0; i < n; i++) du[i] *= t;
                dug *= t;
            }
        }
        memcpy(x_ls, S->x, sizeof(double) * n);
        double step = 1.0;
        for (int64_t i = 0; i < n; i++) S->x[i] = x_ls[i] + step * du[i];
        const double expected = 1e-4 * dug;
        double Ethr = E0 + expected * step;
        int bt = 0;
        for (; bt < 20; bt++) {
            const double E1 = scene_eval(S, S->x, NULL, NULL);
            if (E1 < Ethr) break;
            step *= 0.5;
            for (int64_t i = 0; i < n; i++) S->x[i] = x_ls[i] + step * du[i];
            Ethr = E0 + expected * step;
            st->ls_bt_iterations++;
        }
        if (bt == 20) {
            result = 4; /* TooManyArmijoIterations */
            break;
        }
    }
    st->result = result;
    st->newton_iterations = it;
    free(projected);
    free(bproj);
    free(active);
    free_pattern(&P);
    free(grad);
    free(rhs);
    free(du);
    free(x_ls);
    free(He);
    free(vals);
    return 0;
}

/* Element evaluation + projection + assembly of one Newton iteration (the
 * "elem Hessians/s" unit), timed by the caller. Returns the energy. */
ORACLE_API double oracle_eval_project_assemble(oracle_scene* S, int project, double eps, double* out_vals_or_null)
{
    pattern P;
    build_pattern(S, &P);
    const int64_t n = 3 * (int64_t)S->nv;
    double* grad = (double*)malloc(sizeof(double) * n);
    const int nd = 3 * NPE(S);
    double* He = (double*)malloc(sizeof(double) * nd * nd * (int64_t)(S->ne > 0 ? S->ne : 1));
    double* vals = (double*)malloc(sizeof(double) * 9 * P.nnzb);
    const double E = scene_eval(S, S->x, grad, He);
    if (project) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int e = 0; e < S->ne; e++) oracle_project(He + (int64_t)nd * nd * e, nd, eps, 0);
    }
    assemble(S, &P, He, vals);
    if (out_vals_or_null) memcpy(out_vals_or_null, vals, sizeof(double) * 9 * P.nnzb);
    free_pattern(&P);
    free(grad);
    free(He);
    free(vals);
    return E;
}

/* Adjoint sensitivity of a quasistatic equilibrium x with respect to per-element
 * parameters p_e = log(mu_e), where lambda_e = lambda_over_mu * mu_e (SymX has no
 * adjoint; this composes its pieces: the unprojected Hessian assembled as in
 * ElementHessians.cpp:224-256 and the PCG of solve_pcg.h:83-232).
 *   H lam = dL_dx                 (H unprojected, PCG to rel_tol, no abs_tol)
 *   dL/dp_e = -lam_e^T d^2E_e/dx dp_e
 * Both Lame parameters scale with mu_e, so E_e is linear in mu_e and
 * d^2E_e/dx dp_e = mu_e d(g_e)/d(mu_e) = g_e, the element gradient itself. */
ORACLE_API int oracle_adjoint_logmu(const oracle_scene* S, const double* dL_dx, double rel_tol, int max_iter, double* lam_out, double* dL_dp,
                                    oracle_pcg_info* info)
{
    if (!S->mu_e || NPE(S) != 4) return 1;
    pattern P;
    build_pattern(S, &P);
    const int ne = S->ne;
    double* He = (double*)malloc(sizeof(double) * 144 * (int64_t)(ne > 0 ? ne : 1));
    double* ge = (double*)malloc(sizeof(double) * 12 * (int64_t)(ne > 0 ? ne : 1));
    double* vals = (double*)malloc(sizeof(double) * 9 * P.nnzb);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int e = 0; e < ne; e++) {
        double xe[12], Xe[12], Ee = 0.0;
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) {
                const int64_t v = S->tets[4 * (int64_t)e + a];
                xe[3 * a + d] = S->x[3 * v + d];
                Xe[3 * a + d] = S->X[3 * v + d];
            }
        bad |= oracle_nh_tet4(xe, Xe, S->lambda_over_mu * S->mu_e[e], S->mu_e[e], &Ee, ge + 12 * (int64_t)e, He + 144 * (int64_t)e);
    }
    int rc = bad ? 2 : 0;
    if (!rc) {
        assemble(S, &P, He, vals);
        oracle_pcg_bsr(S->nv, P.rp, P.cols, vals, dL_dx, lam_out, 0.0, rel_tol, max_iter, 1, info);
        for (int e = 0; e < ne; e++) {
            double acc = 0.0;
            for (int a = 0; a < 4; a++)
                for (int d = 0; d < 3; d++) acc += lam_out[3 * (int64_t)S->tets[4 * (int64_t)e + a] + d] * ge[12 * (int64_t)e + 3 * a + d];
            dL_dp[e] = -acc;
        }
    }
    free_pattern(&P);
    free(He);
    free(ge);
    free(vals);
    return rc;
}

ORACLE_API int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
