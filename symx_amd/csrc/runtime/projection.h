// Device code of the PD projection (project_to_PD.cpp:13-32 semantics): eigenvalues
// below eps become eps (clamp) or -lambda (mirror); the matrix is rebuilt as
// V f(Lambda) V^T when some eigenvalue is below eps, and left untouched otherwise.
//
// Eigen-decompositions run as cyclic Jacobi in registers. All loops over matrix
// indices are unrolled at compile time (static_for), so every array index is a
// constant and the arrays stay in VGPRs instead of scratch memory.
//
// Element Hessians of 4-node potentials whose block rows sum to zero (energies
// invariant under rigid translation, e.g. every elastic energy) have the three
// translations as exact null vectors. For those, H = U M U^T with U = W (x) I3
// and W an orthonormal basis of the complement of (1,1,1,1); projecting the 9x9
// M = U^T H U and adding f(0) on the null space gives the same spectral function
// as the 12x12 eigen-decomposition (the null eigenvalues are round-off in the
// 12x12 solve too), at about half the registers and a third of the flops.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace symx_hip
{
    template <int I, int E, typename F>
    __host__ __device__ __forceinline__ void static_for(F&& f)
    {
        if constexpr (I < E) {
            f(std::integral_constant<int, I>());
            static_for<I + 1, E>(f);
        }
    }

    template <int N>
    struct Sym
    {
        static constexpr int size = N * (N + 1) / 2;
        __host__ __device__ static constexpr int idx(int i, int j)
        {
            return (i <= j) ? (i * N - (i * (i - 1)) / 2 + (j - i)) : (j * N - (j * (j - 1)) / 2 + (i - j));
        }
    };

    // One Jacobi rotation zeroing a(P,Q); updates the upper triangle of a and the
    // columns P, Q of v. t solves t^2 + 2 theta t - 1 = 0 with the smaller root,
    // written without dividing by a(P,Q) (t = 0 when a(P,Q) = 0).
    template <int N, int P, int Q>
    __device__ __forceinline__ void jacobi_rotate(double (&a)[Sym<N>::size], double (&v)[N][N])
    {
        constexpr int PP = Sym<N>::idx(P, P), QQ = Sym<N>::idx(Q, Q), PQ = Sym<N>::idx(P, Q);
        const double apq = a[PQ];
        const double app = a[PP];
        const double aqq = a[QQ];
        const double d = aqq - app;
        const double r = sqrt(d * d + 4.0 * apq * apq);
        const double den = d + copysign(r, d);
        const double t = (den != 0.0) ? (2.0 * apq) / den : 0.0;
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double s = t * c;
        a[PP] = app - t * apq;
        a[QQ] = aqq + t * apq;
        a[PQ] = 0.0;
        static_for<0, N>([&](auto Rc) {
            constexpr int R = decltype(Rc)::value;
            if constexpr (R != P && R != Q) {
                constexpr int RP = Sym<N>::idx(R, P), RQ = Sym<N>::idx(R, Q);
                const double g = a[RP];
                const double h = a[RQ];
                a[RP] = c * g - s * h;
                a[RQ] = s * g + c * h;
            }
        });
        static_for<0, N>([&](auto Rc) {
            constexpr int R = decltype(Rc)::value;
            const double g = v[R][P];
            const double h = v[R][Q];
            v[R][P] = c * g - s * h;
            v[R][Q] = s * g + c * h;
        });
        // Keep rotations in order: without this the scheduler defers the (off the
        // critical path) V updates across many rotations and runs out of registers.
        __builtin_amdgcn_sched_barrier(0);
    }

    // a (upper triangle) -> eigenvalues on the diagonal, v -> eigenvectors (columns).
    template <int N>
    __device__ __forceinline__ void jacobi_eig(double (&a)[Sym<N>::size], double (&v)[N][N])
    {
        static_for<0, N>([&](auto Ic) {
            static_for<0, N>([&](auto Jc) { v[decltype(Ic)::value][decltype(Jc)::value] = (decltype(Ic)::value == decltype(Jc)::value) ? 1.0 : 0.0; });
        });
        double fro = 0.0;
        static_for<0, Sym<N>::size>([&](auto Kc) { fro += a[decltype(Kc)::value] * a[decltype(Kc)::value]; });
        const double tol = 1e-32 * fro;
        for (int sweep = 0; sweep < 16; sweep++) {
            double off = 0.0;
            static_for<0, N>([&](auto Pc) {
                static_for<decltype(Pc)::value + 1, N>([&](auto Qc) {
                    const double x = a[Sym<N>::idx(decltype(Pc)::value, decltype(Qc)::value)];
                    off += x * x;
                });
            });
            if (!(off > tol)) break;
            static_for<0, N>([&](auto Pc) {
                static_for<decltype(Pc)::value + 1, N>([&](auto Qc) { jacobi_rotate<N, decltype(Pc)::value, decltype(Qc)::value>(a, v); });
            });
        }
    }

    __device__ __forceinline__ double pd_map(double l, double eps, bool mirror, bool& changed)
    {
        if (l < eps) {
            changed = true;
            return mirror ? -l : eps;
        }
        return l;
    }

    // Full projection of a symmetric N x N matrix given by its upper triangle.
    template <int N>
    __device__ __forceinline__ bool project_sym(double (&a)[Sym<N>::size], double eps, bool mirror)
    {
        double v[N][N];
        jacobi_eig<N>(a, v);
        double lam[N];
        bool changed = false;
        static_for<0, N>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            lam[I] = pd_map(a[Sym<N>::idx(I, I)], eps, mirror, changed);
        });
        if (changed) {
            static_for<0, N>([&](auto Ic) {
                static_for<decltype(Ic)::value, N>([&](auto Jc) {
                    constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                    double acc = 0.0;
                    static_for<0, N>([&](auto Kc) {
                        constexpr int K = decltype(Kc)::value;
                        acc += v[I][K] * lam[K] * v[J][K];
                    });
                    a[Sym<N>::idx(I, J)] = acc;
                });
            });
        }
        return changed;
    }

    // 1/x for |x| >= DBL_MIN: hardware reciprocal estimate and two Newton steps
    // (r <- r + r (1 - x r)), within an ulp or so of the rounded quotient.
    __device__ __forceinline__ double rcp_nr(double x)
    {
        double r = __builtin_amdgcn_rcp(x);
        r = fma(r, fma(-x, r, 1.0), r);
        r = fma(r, fma(-x, r, 1.0), r);
        return r;
    }

    // 1/x with a single Newton step (relative error ~1e-15 on gfx950's estimate): enough for
    // the signs of Sturm pivots.
    __device__ __forceinline__ double rcp_nr1(double x)
    {
        const double r = __builtin_amdgcn_rcp(x);
        return fma(r, fma(-x, r, 1.0), r);
    }

    // Fixed pseudo-random start vector components in (-1, 1) for inverse iteration.
    __host__ __device__ inline double start_component(int j, int i)
    {
        uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + (uint64_t)(j + 1) * 0xBF58476D1CE4E5B9ull;
        h ^= h >> 31;
        h *= 0x94D049BB133111EBull;
        h ^= h >> 29;
        return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }

    // Householder tridiagonalisation in registers (LAPACK dsytd2 arithmetic) of the
    // symmetric N x N matrix a (upper triangle): T = Q^T A Q with diagonal d and
    // off-diagonal e (e[N-1] = 0), Q = H_0 ... H_{N-3}, H_k = I - tau_k v_k v_k^T,
    // v_k(k+1) = 1 and v_k(k+2..N-1) left in a(k, k+2..N-1).
    template <int N>
    __device__ __forceinline__ void tridiag_reg(double (&a)[Sym<N>::size], double (&d)[N], double (&e)[N], double (&tau)[N])
    {
        static_for<0, N - 2>([&](auto Kc) {
            constexpr int K = decltype(Kc)::value;
            double s = 0.0;
            static_for<K + 2, N>([&](auto Ic) {
                const double x = a[Sym<N>::idx(K, decltype(Ic)::value)];
                s += x * x;
            });
            const double x0 = a[Sym<N>::idx(K, K + 1)];
            double beta = x0, t = 0.0, invd = 0.0;
            if (s != 0.0) {
                beta = -copysign(sqrt(x0 * x0 + s), x0);
                t = (beta - x0) / beta;
                invd = 1.0 / (x0 - beta);
            }
            double v[N], p[N];
            v[K + 1] = 1.0;
            static_for<K + 2, N>([&](auto Ic) {
                constexpr int I = decltype(Ic)::value;
                v[I] = a[Sym<N>::idx(K, I)] * invd;
                a[Sym<N>::idx(K, I)] = v[I];
            });
            static_for<K + 1, N>([&](auto Ic) {
                constexpr int I = decltype(Ic)::value;
                double acc = 0.0;
                static_for<K + 1, N>([&](auto Jc) { acc += a[Sym<N>::idx(I, decltype(Jc)::value)] * v[decltype(Jc)::value]; });
                p[I] = t * acc;
            });
            double kk = 0.0;
            static_for<K + 1, N>([&](auto Ic) { kk += p[decltype(Ic)::value] * v[decltype(Ic)::value]; });
            kk *= 0.5 * t;
            static_for<K + 1, N>([&](auto Ic) { p[decltype(Ic)::value] -= kk * v[decltype(Ic)::value]; });
            static_for<K + 1, N>([&](auto Ic) {
                static_for<decltype(Ic)::value, N>([&](auto Jc) {
                    constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                    a[Sym<N>::idx(I, J)] -= v[I] * p[J] + p[I] * v[J];
                });
            });
            d[K] = a[Sym<N>::idx(K, K)];
            e[K] = beta;
            tau[K] = t;
        });
        d[N - 2] = a[Sym<N>::idx(N - 2, N - 2)];
        e[N - 2] = a[Sym<N>::idx(N - 2, N - 1)];
        d[N - 1] = a[Sym<N>::idx(N - 1, N - 1)];
        e[N - 1] = 0.0;
        tau[N - 2] = 0.0;
        tau[N - 1] = 0.0;
    }

    // Number of eigenvalues of T below x0, x1, x2 from the leading principal minors of
    // T - x I, p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2} with p_{-1} = 1: the count is the
    // number of sign changes along p_{-1}, p_0, ..., p_{N-1} (Sturm), the same as the number of
    // negative ratios q_i = p_i / p_{i-1} that sturm3_reg forms with a reciprocal per step;
    // here a step is one FMA deep. An exact zero p_i is replaced by -2^-100 p_{i-1}, i.e. the
    // ratio q_i by -2^-100 (sturm3_reg's -pivmin), which also carries the recurrence across an
    // exact split (e_i = 0). d, e2 (= e^2) and x must be scaled so that ||T|| < 1: |p_i| then
    // stays below (1 + sqrt 2)^N, and the zero replacements (at most N of them) stay far
    // above the underflow threshold.
    template <int N>
    __host__ __device__ __forceinline__ void sturm3_det(const double (&d)[N], const double (&e2)[N], const double (&x)[3], int (&c)[3])
    {
        double p0[3], p1[3];
        static_for<0, 3>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            p0[J] = 1.0;
            p1[J] = d[0] - x[J];
            if (p1[J] == 0.0) p1[J] = -0x1p-100;
            c[J] = p1[J] < 0.0;
        });
        static_for<1, N>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            static_for<0, 3>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                double p2 = fma(d[I] - x[J], p1[J], -e2[I - 1] * p0[J]);
                if (p2 == 0.0) p2 = -0x1p-100 * p1[J];
                c[J] += (p2 < 0.0) != (p1[J] < 0.0);
                p0[J] = p1[J];
                p1[J] = p2;
            });
        });
    }

    // Number of eigenvalues of T below x0, x1, x2 (three interleaved Sturm sequences of
    // T - x I; a pivot below pivmin counts as -pivmin, dlaneg).
    template <int N>
    __device__ __forceinline__ void sturm3_reg(const double (&d)[N], const double (&e)[N], const double (&x)[3], double pivmin,
                                               int (&c)[3])
    {
        double q[3];
        static_for<0, 3>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            q[J] = d[0] - x[J];
            if (fabs(q[J]) < pivmin) q[J] = -pivmin;
            c[J] = q[J] < 0.0;
        });
        static_for<1, N>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            const double e2 = e[I - 1] * e[I - 1];
            static_for<0, 3>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                q[J] = (d[I] - x[J]) - e2 * rcp_nr1(q[J]);  // |q| >= pivmin keeps e2 / q finite
                if (fabs(q[J]) < pivmin) q[J] = -pivmin;
                c[J] += q[J] < 0.0;
            });
        });
    }

    // y <- (T - lam I)^{-1} y: LU with partial pivoting fused with the forward substitution
    // (dlagtf, dlagts), pivots below ptol replaced by +-ptol (bigr = 1 / ptol).
    template <int N>
    __device__ __forceinline__ void tri_solve_reg(const double (&d)[N], const double (&e)[N], double lam, double ptol, double bigr,
                                                  double (&y)[N])
    {
        double u1[N], u2[N], ur[N];
        double a = d[0] - lam, c = e[0], r0 = y[0];
        static_for<0, N - 1>([&](auto Kc) {
            constexpr int K = decltype(Kc)::value;
            const double b = e[K], an = d[K + 1] - lam, cn = (K + 1 < N - 1) ? e[K + 1] : 0.0;
            const double r1 = y[K + 1];
            const bool sw = fabs(a) < fabs(b);  // rows K and K + 1 interchanged
            const double piv = sw ? b : a, oth = sw ? a : b;
            const double inv = fabs(piv) >= 2.2250738585072014e-308 ? rcp_nr(piv) : 0.0;
            const double mult = oth * inv;  // |mult| <= 1; a = b = 0 gives 0
            u1[K] = sw ? an : c;
            u2[K] = sw ? cn : 0.0;
            y[K] = sw ? r1 : r0;
            r0 = sw ? r0 - mult * r1 : r1 - mult * r0;
            const double an2 = sw ? c - mult * an : an - mult * c;
            c = sw ? -mult * cn : cn;
            a = an2;
            ur[K] = fabs(piv) < ptol ? (piv < 0.0 ? -bigr : bigr) : inv;
        });
        ur[N - 1] = fabs(a) < ptol ? (a < 0.0 ? -bigr : bigr) : rcp_nr(a);
        y[N - 1] = r0 * ur[N - 1];
        static_for<0, N - 1>([&](auto Kc) {
            constexpr int K = N - 2 - decltype(Kc)::value;
            double s = y[K] - u1[K] * y[K + 1];
            if constexpr (K + 2 < N) s -= u2[K] * y[K + 2];
            y[K] = s * ur[K];
        });
    }

    // Orthonormal basis of the complement of (1,1,1,1) (Helmert contrasts).
    __device__ __forceinline__ constexpr double helmert(int a, int alpha)
    {
        constexpr double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637, r12 = 0.28867513459481288225;
        return alpha == 0 ? (a == 0 ? r2 : (a == 1 ? -r2 : 0.0))
             : alpha == 1 ? (a <= 1 ? r6 : (a == 2 ? -2.0 * r6 : 0.0))
                          : (a <= 2 ? r12 : -3.0 * r12);
    }

    // Upper-block storage index of block (a, b), a <= b, for 4 blocks per dimension.
    __host__ __device__ constexpr int ub4(int a, int b) { return a * 4 - (a * (a - 1)) / 2 + (b - a); }

    // Entry (3a+i, 3b+j) of a 12x12 element Hessian stored as upper 3x3 blocks.
    template <int A, int B, int I, int J>
    __device__ __forceinline__ double h_entry(const double* h)
    {
        if constexpr (A <= B) return h[ub4(A, B) * 9 + 3 * I + J];
        else return h[ub4(B, A) * 9 + 3 * J + I];
    }
}
