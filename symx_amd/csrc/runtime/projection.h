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

#include <type_traits>

namespace symx_hip
{
    template <int I, int E, typename F>
    __device__ __forceinline__ void static_for(F&& f)
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
