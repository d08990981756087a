// Ahead-of-time kernels of the backend (gfx950): reductions, PD projection,
// deterministic assembly into the sliced block matrix, block-Jacobi PCG and the
// Newton vector operations. Generated element kernels live in hiprtc modules.
//
// Reductions never use atomics or device-wide fences. A reduction kernel runs a
// fixed grid (at most MAX_GROUPS workgroups, each owning a contiguous chunk of
// rows) and writes one partial per workgroup; the scalar is formed by the next
// kernel on the stream, which sums the partials in index order (every workgroup
// of a PCG kernel does this redundantly, so all of them see the same alpha/beta
// without synchronising). Results are therefore bitwise repeatable for a given
// problem size.
//
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// xcd_remap() gives every XCD a contiguous range of logical workgroups, so rows
// that share neighbours (gathers through the mesh connectivity) hit the same L2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.h"
#include "projection.h"

namespace symx_hip
{
    constexpr int NT = 256;  // threads per workgroup
    constexpr int XCDS = 8;
    constexpr int MAX_GROUPS = 1024;

    int reduction_groups(int64_t n)
    {
        int64_t g = (n + NT - 1) / NT;
        if (g < 1) g = 1;
        if (g > MAX_GROUPS) g = MAX_GROUPS;
        return (int)g;
    }

    static inline unsigned grid_for(int64_t n) { return (unsigned)((n + NT - 1) / NT > 0 ? (n + NT - 1) / NT : 1); }

    // Logical workgroup index: XCD x (hardware workgroups b with b % 8 == x) gets the
    // contiguous logical range [x*per + min(x, extra), ...).
    __device__ __forceinline__ int64_t xcd_remap()
    {
        const int64_t b = blockIdx.x, nb = gridDim.x;
        const int64_t x = b % XCDS, i = b / XCDS;
        const int64_t per = nb / XCDS, extra = nb % XCDS;
        return x * per + (x < extra ? x : extra) + i;
    }

    // Contiguous row chunk [r0, r1) of logical group w out of G (chunks are multiples of 64 rows).
    __device__ __forceinline__ void row_chunk(int64_t n, int64_t w, int64_t G, int64_t& r0, int64_t& r1)
    {
        const int64_t per = (((n + G - 1) / G) + 63) / 64 * 64;
        r0 = w * per < n ? w * per : n;
        r1 = r0 + per < n ? r0 + per : n;
    }

    __device__ inline double wave_sum(double v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    }
    // max that keeps NaN visible (a NaN residual must not look converged)
    __device__ inline double nanmax(double a, double b) { return (a != a || b != b) ? (a + b) : (a > b ? a : b); }
    __device__ inline double wave_max(double v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, 64));
        return v;
    }

    // Workgroup sum in a fixed order, returned to every thread.
    __device__ inline double block_sum(double v, double* lds)
    {
        v = wave_sum(v);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) lds[w] = v;
        __syncthreads();
        double r = 0.0;
        for (int i = 0; i < NT / 64; i++) r += lds[i];
        return r;
    }
    __device__ inline double block_max(double v, double* lds)
    {
        v = wave_max(v);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) lds[w] = v;
        __syncthreads();
        double r = lds[0];
        for (int i = 1; i < NT / 64; i++) r = nanmax(r, lds[i]);
        return r;
    }

    // Fixed-order sum of G partials p[i*stride + off], returned to every thread.
    __device__ inline double sum_partials(const double* __restrict__ p, int64_t G, int stride, int off, double* lds)
    {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < G; i += NT) acc += p[i * stride + off];
        return block_sum(acc, lds);
    }

    // ---- plain sums ----------------------------------------------------------------
    __global__ __launch_bounds__(NT) void k_sum(const double* __restrict__ in, int64_t n, int stride, int off, double* out)
    {
        __shared__ double lds[NT / 64];
        const double r = sum_partials(in, n, stride, off, lds);
        if (threadIdx.x == 0) *out = r;
    }
    void launch_sum(hipStream_t s, const double* d_in, int64_t n, double* d_out)
    {
        k_sum<<<1, NT, 0, s>>>(d_in, n, 1, 0, d_out);
    }

    // ---- gradient assembly: grad[3b+i] = sum of element gradient blocks -----------
    __global__ __launch_bounds__(NT) void k_grad_assemble(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                          const double* __restrict__ G, double* __restrict__ grad, int32_t nb)
    {
        const int64_t b = xcd_remap() * NT + threadIdx.x;
        if (b >= nb) return;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        const uint32_t k1 = off[b + 1];
        for (uint32_t k = off[b]; k < k1; k++) {
            const double* src = G + (int64_t)list[k] * 3;
            g0 += src[0];
            g1 += src[1];
            g2 += src[2];
        }
        grad[3 * b + 0] = g0;
        grad[3 * b + 1] = g1;
        grad[3 * b + 2] = g2;
    }
    void launch_grad_assemble(hipStream_t s, const uint32_t* off, const uint32_t* list, const double* G, double* grad, int32_t nb)
    {
        if (nb <= 0) return;
        k_grad_assemble<<<grid_for(nb), NT, 0, s>>>(off, list, G, grad, nb);
    }

    // ---- PD projection (device code in projection.h) ---------------------------------
    // Full-size projection of an element Hessian stored as upper 3x3 blocks.
    template <int NB>
    __device__ __forceinline__ bool project_element_full(double* __restrict__ h, double eps, bool mirror)
    {
        constexpr int N = 3 * NB;
        double a[Sym<N>::size];
        static_for<0, NB>([&](auto Bi) {
            static_for<decltype(Bi)::value, NB>([&](auto Bj) {
                constexpr int bi = decltype(Bi)::value, bj = decltype(Bj)::value;
                constexpr int ub = bi * NB - (bi * (bi - 1)) / 2 + (bj - bi);
                static_for<0, 3>([&](auto Ic) {
                    static_for<0, 3>([&](auto Jc) {
                        constexpr int i = decltype(Ic)::value, j = decltype(Jc)::value;
                        if constexpr (bi != bj || i <= j) a[Sym<N>::idx(3 * bi + i, 3 * bj + j)] = h[ub * 9 + 3 * i + j];
                    });
                });
            });
        });
        const bool changed = project_sym<N>(a, eps, mirror);
        if (changed) {
            static_for<0, NB>([&](auto Bi) {
                static_for<decltype(Bi)::value, NB>([&](auto Bj) {
                    constexpr int bi = decltype(Bi)::value, bj = decltype(Bj)::value;
                    constexpr int ub = bi * NB - (bi * (bi - 1)) / 2 + (bj - bi);
                    static_for<0, 3>([&](auto Ic) {
                        static_for<0, 3>([&](auto Jc) {
                            constexpr int i = decltype(Ic)::value, j = decltype(Jc)::value;
                            h[ub * 9 + 3 * i + j] = a[Sym<N>::idx(3 * bi + i, 3 * bj + j)];
                        });
                    });
                });
            });
        }
        return changed;
    }

    // 4-node element Hessian with the rigid translations as null vectors: project the
    // 9x9 restriction M = U^T H U (U = W (x) I3) and map the null space with f(0).
    // Returns false (and leaves h untouched) if the block rows do not sum to zero.
    __device__ __forceinline__ bool project_element_reduced4(double* __restrict__ h, double eps, bool mirror, bool& changed)
    {
        double scale = 0.0, rowsum = 0.0;
        static_for<0, 90>([&](auto Kc) { scale = fmax(scale, fabs(h[decltype(Kc)::value])); });
        static_for<0, 4>([&](auto Ac) {
            static_for<0, 3>([&](auto Ic) {
                static_for<0, 3>([&](auto Jc) {
                    constexpr int A = decltype(Ac)::value, I = decltype(Ic)::value, J = decltype(Jc)::value;
                    const double s = h_entry<A, 0, I, J>(h) + h_entry<A, 1, I, J>(h) + h_entry<A, 2, I, J>(h) + h_entry<A, 3, I, J>(h);
                    rowsum = fmax(rowsum, fabs(s));
                });
            });
        });
        if (!(rowsum <= 1e-10 * scale)) return false;
        double m[Sym<9>::size];
        static_for<0, 3>([&](auto Al) {
            static_for<decltype(Al)::value, 3>([&](auto Be) {
                static_for<0, 3>([&](auto Ic) {
                    static_for<0, 3>([&](auto Jc) {
                        constexpr int al = decltype(Al)::value, be = decltype(Be)::value;
                        constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                        if constexpr (al != be || I <= J) {
                            double acc = 0.0;
                            static_for<0, 4>([&](auto Ac) {
                                static_for<0, 4>([&](auto Bc) {
                                    constexpr int A = decltype(Ac)::value, B = decltype(Bc)::value;
                                    constexpr double w = helmert(A, al) * helmert(B, be);
                                    if constexpr (w != 0.0) acc += w * h_entry<A, B, I, J>(h);
                                });
                            });
                            m[Sym<9>::idx(3 * al + I, 3 * be + J)] = acc;
                        }
                    });
                });
            });
        });
        double v[9][9];
        jacobi_eig<9>(m, v);
        double lam[9];
        bool any = false;
        static_for<0, 9>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            lam[I] = pd_map(m[Sym<9>::idx(I, I)], eps, mirror, any);
        });
        // Mp = V f(Lambda) V^T (always rebuilt: the null eigenvalues are below eps)
        static_for<0, 9>([&](auto Ic) {
            static_for<decltype(Ic)::value, 9>([&](auto Jc) {
                constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                double acc = 0.0;
                static_for<0, 9>([&](auto Kc) {
                    constexpr int K = decltype(Kc)::value;
                    acc += v[I][K] * lam[K] * v[J][K];
                });
                m[Sym<9>::idx(I, J)] = acc;
            });
        });
        const double null_term = mirror ? 0.0 : 0.25 * eps;  // f(0) on span{translations}: eps T T^T, T T^T = (1/4) 1 1^T (x) I3
        static_for<0, 4>([&](auto Ac) {
            static_for<decltype(Ac)::value, 4>([&](auto Bc) {
                static_for<0, 3>([&](auto Ic) {
                    static_for<0, 3>([&](auto Jc) {
                        constexpr int A = decltype(Ac)::value, B = decltype(Bc)::value;
                        constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                        double acc = (I == J) ? null_term : 0.0;
                        static_for<0, 3>([&](auto Al) {
                            static_for<0, 3>([&](auto Be) {
                                constexpr int al = decltype(Al)::value, be = decltype(Be)::value;
                                constexpr double w = helmert(A, al) * helmert(B, be);
                                if constexpr (w != 0.0) acc += w * m[Sym<9>::idx(3 * al + I, 3 * be + J)];
                            });
                        });
                        h[ub4(A, B) * 9 + 3 * I + J] = acc;
                    });
                });
            });
        });
        changed = true;
        return true;
    }

    __device__ inline void count_changed(bool changed, unsigned long long* count)
    {
        if (count != nullptr) {
            const unsigned long long m = __ballot(changed);
            if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(count, (unsigned long long)__popcll(m));
        }
    }

    // 4-block elements: reduced 9x9 path; elements that are not translation invariant
    // are flagged for k_project_ub_full.
    __global__ __launch_bounds__(64) void k_project_ub4_reduced(double* __restrict__ H, int64_t n_mats, double eps, int mirror,
                                                               unsigned long long* count, uint8_t* __restrict__ needs_full,
                                                               const uint8_t* __restrict__ select)
    {
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        bool changed = false;
        if (e < n_mats) {
            if (select == nullptr || select[e]) {
                const bool ok = project_element_reduced4(H + e * 90, eps, mirror != 0, changed);
                needs_full[e] = ok ? 0 : 1;
            }
            else {
                needs_full[e] = 0;
            }
        }
        count_changed(changed, count);
    }

    template <int NB>
    __global__ __launch_bounds__(64) void k_project_ub_full(double* __restrict__ H, int64_t n_mats, double eps, int mirror,
                                                           unsigned long long* count, const uint8_t* __restrict__ only)
    {
        constexpr int NUB = NB * (NB + 1) / 2;
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        bool changed = false;
        if (e < n_mats && (only == nullptr || only[e])) changed = project_element_full<NB>(H + e * (NUB * 9), eps, mirror != 0);
        count_changed(changed, count);
    }

    // Large element Hessians (8 and more DoF blocks: Hex8, Tet10, Hex27): one workgroup
    // of 512 threads per matrix, the matrix in LDS.
    //  1. Fast path for translation-invariant elements (every block row sums to zero, as
    //     for every elastic energy): the translations are exact null vectors, so shifting
    //     them up by c (adding c/NB to the diagonal entries of every 3x3 block) leaves the
    //     other eigenvalues alone. If the Cholesky factorisation of the shifted matrix minus
    //     eps I succeeds, the only eigenvalues below eps are the three zeros, and the
    //     projection adds f(0) on the translation space: H + eps/NB (1 1^T (x) I3) (clamp) or
    //     H unchanged (mirror), which the register kernels of 4-node elements also do.
    //  2. Otherwise parallel cyclic Jacobi with the round-robin pairing of Brent and Luk
    //     (n/2 disjoint rotations per round, n-1 rounds per sweep): the rotation of
    //     jacobi_rotate, A <- J^T A J as a column pass then a row pass with the rotated
    //     2x2 block set exactly, V <- V J, then V f(Lambda) V^T when some eigenvalue < eps.
    template <int NB>
    __global__ __launch_bounds__(512) void k_project_ub_lds(double* __restrict__ H, int64_t n_mats, double eps, int mirror,
                                                           unsigned long long* count, const uint8_t* __restrict__ select)
    {
        constexpr int NT = 512, N = 3 * NB, NP = (N + 1) & ~1, LD = NP + 1, NUB = NB * (NB + 1) / 2, M = NP / 2;
        __shared__ double A[NP * LD];
        __shared__ double V[NP * LD];
        __shared__ double c_[M], s_[M], app_[M], aqq_[M];
        __shared__ int pq[2 * M];
        __shared__ double red[NT / 64];
        __shared__ double lam[NP];
        __shared__ int flag;
        const int64_t e = blockIdx.x;
        if (e >= n_mats) return;
        if (select != nullptr && !select[e]) return;
        const int tid = threadIdx.x;
        double* h = H + e * (int64_t)(NUB * 9);
        auto ubi = [](int a, int b) { return a * NB - (a * (a - 1)) / 2 + (b - a); };
        for (int t = tid; t < NP * NP; t += NT) {
            const int r = t / NP, c = t - r * NP;
            double v = 0.0;
            if (r < N && c < N) {
                const int a = r / 3, i = r - 3 * a, b = c / 3, j = c - 3 * b;
                v = a <= b ? h[ubi(a, b) * 9 + 3 * i + j] : h[ubi(b, a) * 9 + 3 * j + i];
            }
            A[r * LD + c] = v;
        }
        __syncthreads();
        auto block_max = [&](double v) {
            for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
            if ((tid & 63) == 0) red[tid >> 6] = v;
            __syncthreads();
            double r = red[0];
            for (int w = 1; w < NT / 64; w++) r = fmax(r, red[w]);
            __syncthreads();
            return r;
        };
        auto block_sum = [&](double v) {
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((tid & 63) == 0) red[tid >> 6] = v;
            __syncthreads();
            double r = 0.0;
            for (int w = 0; w < NT / 64; w++) r += red[w];
            __syncthreads();
            return r;
        };
        // ---- 1. translation-invariant fast path --------------------------------------
        double amax = 0.0, rsum = 0.0;
        for (int t = tid; t < N * N; t += NT) amax = fmax(amax, fabs(A[(t / N) * LD + t % N]));
        for (int t = tid; t < N * 3; t += NT) {  // row r, column offset j: sum over blocks b of A[r][3b + j]
            const int r = t / 3, j = t - 3 * r;
            double acc = 0.0;
            for (int b = 0; b < NB; b++) acc += A[r * LD + 3 * b + j];
            rsum = fmax(rsum, fabs(acc));
        }
        const double scale = block_max(amax);
        const double row_sum = block_max(rsum);
        if (row_sum <= 1e-10 * scale * NB) {
            const double shift = scale * NB;
            for (int t = tid; t < N * N; t += NT) {
                const int r = t / N, c = t - r * N;
                V[r * LD + c] = A[r * LD + c] + ((r % 3) == (c % 3) ? shift / NB : 0.0) - (r == c ? eps : 0.0);
            }
            if (tid == 0) flag = 0;
            __syncthreads();
            for (int k = 0; k < N; k++) {  // right-looking Cholesky, lower triangle
                if (tid == 0) {
                    const double d = V[k * LD + k];
                    if (!(d > 0.0)) flag = 1;
                    else V[k * LD + k] = sqrt(d);
                }
                __syncthreads();
                if (flag) break;
                const double piv = V[k * LD + k];
                for (int i = k + 1 + tid; i < N; i += NT) V[i * LD + k] /= piv;
                __syncthreads();
                const int m = N - k - 1;
                for (int t = tid; t < m * m; t += NT) {
                    const int i = k + 1 + t / m, j = k + 1 + t % m;
                    if (j <= i) V[i * LD + j] -= V[i * LD + k] * V[j * LD + k];
                }
                __syncthreads();
            }
            if (!flag) {
                if (!mirror) {
                    for (int t = tid; t < NUB * 9; t += NT) {
                        const int ij = t % 9;
                        if (ij / 3 == ij % 3) h[t] += eps / NB;
                    }
                }
                if (tid == 0 && count != nullptr) atomicAdd(count, 1ull);
                return;
            }
        }
        // ---- 2. parallel cyclic Jacobi ----------------------------------------------------
        for (int t = tid; t < NP * NP; t += NT) {
            const int r = t / NP, c = t - r * NP;
            V[r * LD + c] = r == c ? 1.0 : 0.0;
        }
        double fro = 0.0;
        for (int t = tid; t < N * N; t += NT) {
            const int r = t / N, c = t - r * N;
            if (r <= c) fro += A[r * LD + c] * A[r * LD + c];
        }
        const double tol = 1e-32 * block_sum(fro);
        for (int sweep = 0; sweep < 24; sweep++) {
            double off = 0.0;
            for (int t = tid; t < N * N; t += NT) {
                const int r = t / N, c = t - r * N;
                if (r < c) off += A[r * LD + c] * A[r * LD + c];
            }
            if (!(block_sum(off) > tol)) break;
            for (int round = 0; round < NP - 1; round++) {
                if (tid < M) {
                    // circle method: index NP-1 stays, the others rotate
                    int p, q;
                    if (tid == 0) {
                        p = round;
                        q = NP - 1;
                    }
                    else {
                        p = (round + tid) % (NP - 1);
                        q = (round - tid + NP - 1) % (NP - 1);
                    }
                    if (p > q) {
                        const int x = p;
                        p = q;
                        q = x;
                    }
                    const double apq = A[p * LD + q], app = A[p * LD + p], aqq = A[q * LD + q];
                    const double d = aqq - app;
                    const double rr = sqrt(d * d + 4.0 * apq * apq);
                    const double den = d + copysign(rr, d);
                    const double tt = (den != 0.0) ? (2.0 * apq) / den : 0.0;
                    const double c = 1.0 / sqrt(1.0 + tt * tt);
                    c_[tid] = c;
                    s_[tid] = tt * c;
                    app_[tid] = app - tt * apq;
                    aqq_[tid] = aqq + tt * apq;
                    pq[2 * tid] = p;
                    pq[2 * tid + 1] = q;
                }
                __syncthreads();
                for (int t = tid; t < NP * M; t += NT) {  // columns of A and V
                    const int r = t / M, k = t - r * M;
                    const int p = pq[2 * k], q = pq[2 * k + 1];
                    const double c = c_[k], s = s_[k];
                    const double g = A[r * LD + p], hh = A[r * LD + q];
                    A[r * LD + p] = c * g - s * hh;
                    A[r * LD + q] = s * g + c * hh;
                    const double gv = V[r * LD + p], hv = V[r * LD + q];
                    V[r * LD + p] = c * gv - s * hv;
                    V[r * LD + q] = s * gv + c * hv;
                }
                __syncthreads();
                for (int t = tid; t < M * NP; t += NT) {  // rows of A; the rotated 2x2 block exactly
                    const int k = t / NP, col = t - k * NP;
                    const int p = pq[2 * k], q = pq[2 * k + 1];
                    if (col == p) {
                        A[p * LD + p] = app_[k];
                        A[q * LD + p] = 0.0;
                    }
                    else if (col == q) {
                        A[p * LD + q] = 0.0;
                        A[q * LD + q] = aqq_[k];
                    }
                    else {
                        const double c = c_[k], s = s_[k];
                        const double g = A[p * LD + col], hh = A[q * LD + col];
                        A[p * LD + col] = c * g - s * hh;
                        A[q * LD + col] = s * g + c * hh;
                    }
                }
                __syncthreads();
            }
        }
        if (tid == 0) flag = 0;
        __syncthreads();
        for (int t = tid; t < N; t += NT) {
            const double l = A[t * LD + t];
            if (l < eps) {
                lam[t] = mirror ? -l : eps;
                flag = 1;
            }
            else {
                lam[t] = l;
            }
        }
        __syncthreads();
        if (!flag) return;
        for (int t = tid; t < NUB * 9; t += NT) {
            const int u = t / 9, ij = t - 9 * u;
            int a = 0, rem = u;
            while (rem >= NB - a) {
                rem -= NB - a;
                a++;
            }
            const int b = a + rem;
            const int r = 3 * a + ij / 3, c = 3 * b + ij % 3;
            double acc = 0.0;
            for (int k = 0; k < N; k++) acc += V[r * LD + k] * lam[k] * V[c * LD + k];
            h[t] = acc;
        }
        if (tid == 0 && count != nullptr) atomicAdd(count, 1ull);
    }

    int launch_project_ub(hipStream_t s, double* H, int64_t n_mats, int32_t nb, double eps, int32_t mirror, uint8_t* flags,
                          const uint8_t* select, unsigned long long* d_count)
    {
        if (n_mats <= 0) return 0;
        const unsigned g = (unsigned)((n_mats + 63) / 64);
        switch (nb) {
        case 1: k_project_ub_full<1><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        case 2: k_project_ub_full<2><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        case 3: k_project_ub_full<3><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        case 4:
            k_project_ub4_reduced<<<g, 64, 0, s>>>(H, n_mats, eps, mirror, d_count, flags, select);
            k_project_ub_full<4><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, d_count, flags);
            return 0;
        case 8: k_project_ub_lds<8><<<(unsigned)n_mats, 512, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        case 10: k_project_ub_lds<10><<<(unsigned)n_mats, 512, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        case 27: k_project_ub_lds<27><<<(unsigned)n_mats, 512, 0, s>>>(H, n_mats, eps, mirror, d_count, select); return 0;
        default: return -1;
        }
    }

    // ---- progressive projection (NewtonsMethod.cpp:306-330, ElementHessians.cpp:79-165) ----
    __global__ __launch_bounds__(NT) void k_ppn_select(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                      const double* __restrict__ grad, int32_t nblocks, double threshold, PotRanges pr,
                                                      const uint8_t* __restrict__ projected, uint8_t* __restrict__ select,
                                                      unsigned long long* n_active)
    {
        const int64_t b = (int64_t)blockIdx.x * NT + threadIdx.x;
        bool active = false;
        if (b < nblocks) {
            const double m = fmax(fabs(grad[3 * b]), fmax(fabs(grad[3 * b + 1]), fabs(grad[3 * b + 2])));
            active = m >= threshold;
            if (active) {
                for (uint32_t k = off[b]; k < off[b + 1]; k++) {
                    const int64_t gi = list[k];
                    for (int q = 0; q < pr.n; q++) {
                        if (gi >= pr.g_base[q] && gi < pr.g_end[q]) {
                            const int64_t le = (gi - pr.g_base[q]) / pr.nb[q];
                            const int64_t e = pr.e_base[q] + le;
                            if (!projected[e] && (pr.active[q] == nullptr || pr.active[q][le])) select[e] = 1;
                            break;
                        }
                    }
                }
            }
        }
        const unsigned long long m = __ballot(active);
        if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(n_active, (unsigned long long)__popcll(m));
    }
    void launch_ppn_select(hipStream_t s, const uint32_t* off, const uint32_t* list, const double* grad, int32_t nblocks, double threshold,
                           const PotRanges& pr, const uint8_t* projected, uint8_t* select, unsigned long long* n_active)
    {
        if (nblocks <= 0) return;
        k_ppn_select<<<(unsigned)((nblocks + NT - 1) / NT), NT, 0, s>>>(off, list, grad, nblocks, threshold, pr, projected, select, n_active);
    }

    __global__ __launch_bounds__(NT) void k_ppn_commit(uint8_t* __restrict__ projected, uint8_t* __restrict__ select, int64_t n,
                                                      unsigned long long* n_new)
    {
        const int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x;
        bool fresh = false;
        if (e < n) {
            fresh = select[e] != 0;
            if (fresh) projected[e] = 1;
            select[e] = 0;
        }
        const unsigned long long m = __ballot(fresh);
        if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(n_new, (unsigned long long)__popcll(m));
    }
    void launch_ppn_commit(hipStream_t s, uint8_t* projected, uint8_t* select, int64_t n, unsigned long long* n_new)
    {
        if (n <= 0) return;
        k_ppn_commit<<<(unsigned)((n + NT - 1) / NT), NT, 0, s>>>(projected, select, n, n_new);
    }

    // Non-negative doubles order like their bit patterns, so the minimum is an integer atomicMin.
    __global__ __launch_bounds__(NT) void k_plane_max_step(const int32_t* __restrict__ blocks, int32_t n, const double* __restrict__ x,
                                                          const double* __restrict__ du, int32_t axis, double plane, double eta,
                                                          unsigned long long* out)
    {
        const int32_t i = blockIdx.x * NT + threadIdx.x;
        double t = 1.0;
        if (i < n) {
            const int64_t k = 3 * (int64_t)blocks[i] + axis;
            const double d = du[k];
            if (d < 0.0) t = fmin(1.0, fmax(0.0, eta * (x[k] - plane) / -d));
        }
        for (int off = 32; off > 0; off >>= 1) t = fmin(t, __shfl_xor(t, off, 64));
        if ((threadIdx.x & 63) == 0 && t < 1.0) atomicMin(out, (unsigned long long)__double_as_longlong(t));
    }
    void launch_plane_max_step(hipStream_t s, const int32_t* blocks, int32_t n, const double* x, const double* du, int32_t axis, double plane,
                               double eta, unsigned long long* out)
    {
        if (n <= 0) return;
        k_plane_max_step<<<(unsigned)((n + NT - 1) / NT), NT, 0, s>>>(blocks, n, x, du, axis, plane, eta, out);
    }

    __global__ __launch_bounds__(NT) void k_fill_u8(uint8_t* __restrict__ p, int64_t n)
    {
        const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (i < n) p[i] = 0;
    }
    void launch_fill_u8(hipStream_t s, uint8_t* p, int64_t n)
    {
        if (n <= 0) return;
        k_fill_u8<<<(unsigned)((n + NT - 1) / NT), NT, 0, s>>>(p, n);
    }

    template <int N>
    __global__ __launch_bounds__(64) void k_project_dense(double* __restrict__ M, int64_t n_mats, double eps, int mirror,
                                                         uint8_t* changed_out)
    {
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        if (e >= n_mats) return;
        double* m = M + e * (N * N);
        double a[Sym<N>::size];
        static_for<0, N>([&](auto Ic) {
            static_for<decltype(Ic)::value, N>([&](auto Jc) {
                constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                a[Sym<N>::idx(I, J)] = m[I * N + J];
            });
        });
        const bool changed = project_sym<N>(a, eps, mirror != 0);
        if (changed) {
            static_for<0, N>([&](auto Ic) {
                static_for<0, N>([&](auto Jc) {
                    constexpr int I = decltype(Ic)::value, J = decltype(Jc)::value;
                    m[I * N + J] = a[Sym<N>::idx(I, J)];
                });
            });
        }
        if (changed_out) changed_out[e] = changed ? 1 : 0;
    }

    int launch_project_dense(hipStream_t s, double* mats, int64_t n_mats, int32_t n, double eps, int32_t mirror, uint8_t* changed)
    {
        if (n_mats <= 0) return 0;
        const unsigned g = (unsigned)((n_mats + 63) / 64);
        switch (n) {
        case 3: k_project_dense<3><<<g, 64, 0, s>>>(mats, n_mats, eps, mirror, changed); return 0;
        case 6: k_project_dense<6><<<g, 64, 0, s>>>(mats, n_mats, eps, mirror, changed); return 0;
        case 9: k_project_dense<9><<<g, 64, 0, s>>>(mats, n_mats, eps, mirror, changed); return 0;
        case 12: k_project_dense<12><<<g, 64, 0, s>>>(mats, n_mats, eps, mirror, changed); return 0;
        default: return -1;
        }
    }

    // ---- assembly: one thread per block-matrix slot, contributions in fixed order ----
    template <typename TS>
    __global__ __launch_bounds__(NT) void k_bsr_assemble(int64_t n_slots, const uint32_t* __restrict__ coff,
                                                         const uint32_t* __restrict__ clist, const double* __restrict__ H,
                                                         TS* __restrict__ val)
    {
        const int64_t t = xcd_remap() * NT + threadIdx.x;
        if (t >= n_slots) return;
        double acc[9];
#pragma unroll
        for (int k = 0; k < 9; k++) acc[k] = 0.0;
        const uint32_t k1 = coff[t + 1];
        for (uint32_t k = coff[t]; k < k1; k++) {
            const uint32_t code = clist[k];
            const double* h = H + (int64_t)(code >> 1) * 9;
            if (code & 1u) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
#pragma unroll
                    for (int j = 0; j < 3; j++) acc[3 * i + j] += h[3 * j + i];
                }
            }
            else {
#pragma unroll
                for (int k2 = 0; k2 < 9; k2++) acc[k2] += h[k2];
            }
        }
        TS* dst = val + ((t >> 6) * 9) * 64 + (t & 63);
#pragma unroll
        for (int k = 0; k < 9; k++) dst[k * 64] = (TS)acc[k];
    }

    void launch_bsr_assemble(hipStream_t s, const SellPattern& P, const uint32_t* coff, const uint32_t* clist, const double* H,
                             double* v64, float* v32)
    {
        if (P.n_slots <= 0) return;
        if (v64) k_bsr_assemble<double><<<grid_for(P.n_slots), NT, 0, s>>>(P.n_slots, coff, clist, H, v64);
        if (v32) k_bsr_assemble<float><<<grid_for(P.n_slots), NT, 0, s>>>(P.n_slots, coff, clist, H, v32);
    }

    // Closed-form inverse of the symmetric diagonal block of each row, computed in
    // storage precision like BlockedSparseMatrix.h:1198-1214. Layout: dinv[k*n_rows+row].
    template <typename TS>
    __global__ __launch_bounds__(NT) void k_dinv(int32_t n_rows, const int64_t* __restrict__ slice_ptr, const TS* __restrict__ val,
                                                 TS* __restrict__ dinv)
    {
        const int64_t row = xcd_remap() * NT + threadIdx.x;
        if (row >= n_rows) return;
        const int64_t base = slice_ptr[row >> 6];
        const TS* m = val + (base >> 6) * 9 * 64 + (row & 63);  // j = 0 is the diagonal block
        const TS m00 = m[0], m01 = m[64], m02 = m[128], m11 = m[4 * 64], m12 = m[5 * 64], m22 = m[8 * 64];
        const TS c00 = m11 * m22 - m12 * m12;
        const TS c01 = m02 * m12 - m01 * m22;
        const TS c02 = m01 * m12 - m11 * m02;
        const TS det = m00 * c00 + m01 * c01 + m02 * c02;
        TS inv[9];
        if (det != (TS)0) {
            const TS id = (TS)(1.0 / (double)det);
            inv[0] = c00 * id;
            inv[1] = c01 * id;
            inv[2] = c02 * id;
            inv[4] = (m00 * m22 - m02 * m02) * id;
            inv[5] = (m01 * m02 - m00 * m12) * id;
            inv[8] = (m00 * m11 - m01 * m01) * id;
        }
        else {
            inv[0] = inv[1] = inv[2] = inv[4] = inv[5] = inv[8] = (TS)0;
        }
        inv[3] = inv[1];
        inv[6] = inv[2];
        inv[7] = inv[5];
#pragma unroll
        for (int k = 0; k < 9; k++) dinv[(int64_t)k * n_rows + row] = inv[k];
    }

    void launch_dinv(hipStream_t s, const SellPattern& P, const double* v64, const float* v32, double* dinv64, float* dinv32)
    {
        if (P.n_rows <= 0) return;
        if (v64 && dinv64) k_dinv<double><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, P.d_slice_ptr, v64, dinv64);
        if (v32 && dinv32) k_dinv<float><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, P.d_slice_ptr, v32, dinv32);
    }

    // ---- PCG -------------------------------------------------------------------------
    // Iteration k (1-based) = two kernels:
    //   S_k  prologue: rr, rz of iteration k-1 from the partials of U_{k-1} (or init);
    //                  convergence test of solve_pcg.h:200-210 for iteration k-1 (for
    //                  k = 1 the early exits of :116-154); beta = rz / rz_prev.
    //        body:     p_new = z + beta p_old (own row, and on the fly for neighbours),
    //                  q = A p_new, partial p.q.
    //   U_k  prologue: pAp from S_k's partials, alpha = rz / pAp, indefiniteness stop.
    //        body:     x += alpha p, r -= alpha q, z = D^-1 r, partial r.r and r.z.
    // Workgroup 0 publishes the solver state; a finished solve turns every later
    // kernel into a no-op, so the host enqueues batches of iterations.
    // Partials: pa[parity][2*G] (r.r, r.z) written by init / U_k with parity k & 1,
    // pb[G] (p.q) written by S_k.

    __device__ __forceinline__ int64_t pcg_groups(int64_t n_rows) { return n_rows <= 0 ? 1 : ((n_rows + NT - 1) / NT < MAX_GROUPS ? (n_rows + NT - 1) / NT : MAX_GROUPS); }

    template <typename T>
    __device__ __forceinline__ void apply_dinv(const T* __restrict__ dinv, int64_t n_rows, int64_t row, const T (&rv)[3], T (&zv)[3])
    {
#pragma unroll
        for (int i = 0; i < 3; i++)
            zv[i] = dinv[(3 * i + 0) * n_rows + row] * rv[0] + dinv[(3 * i + 1) * n_rows + row] * rv[1] + dinv[(3 * i + 2) * n_rows + row] * rv[2];
    }

    template <typename T>
    __global__ __launch_bounds__(NT) void k_pcg_init(int32_t n_rows, const T* __restrict__ dinv, const T* __restrict__ b,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, T* __restrict__ p0,
                                                     double* __restrict__ pa0)
    {
        __shared__ double lds[NT / 64];
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, gridDim.x, r0, r1);
        double bb = 0.0, rz = 0.0;
        for (int64_t row = r0 + threadIdx.x; row < r1; row += NT) {
            T rv[3], zv[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int64_t k = 3 * row + i;
                rv[i] = b[k];
                x[k] = (T)0;
                r[k] = rv[i];
                p0[k] = (T)0;
            }
            apply_dinv(dinv, n_rows, row, rv, zv);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                z[3 * row + i] = zv[i];
                bb += (double)rv[i] * (double)rv[i];
                rz += (double)rv[i] * (double)zv[i];
            }
        }
        const double sbb = block_sum(bb, lds);
        const double srz = block_sum(rz, lds);
        if (threadIdx.x == 0) {
            pa0[2 * w + 0] = sbb;
            pa0[2 * w + 1] = srz;
        }
    }

    template <typename TS, typename TV>
    __global__ __launch_bounds__(NT) void k_pcg_spmv(int32_t n_rows, int32_t k, const int64_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col, const TS* __restrict__ val,
                                                     const TV* __restrict__ z, const TV* __restrict__ p_old, TV* __restrict__ p_new,
                                                     TV* __restrict__ q, PcgState* st, const double* __restrict__ pa_prev,
                                                     const double* __restrict__ pa_prev2, double* __restrict__ pb)
    {
        __shared__ double lds[NT / 64];
        if (st->done) return;
        const int64_t G = gridDim.x;
        const double rr = sum_partials(pa_prev, G, 2, 0, lds);
        const double rz = sum_partials(pa_prev, G, 2, 1, lds);
        double beta = 0.0;
        int done = 0, converged = 0, it = k - 1;
        double err = 0.0;
        if (k == 1) {
            const double bb = rr;  // x = 0, so r = b
            if (bb < st->abs_tol * st->abs_tol) {
                done = 1;
                converged = 1;
            }
            else {
                err = sqrt(rr / bb);
                if (err < st->abs_tol) {
                    done = 1;
                    converged = 1;
                }
                else if (st->max_iter <= 0) {
                    done = 1;
                }
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                st->bb = bb;
                st->err0 = err;
                st->rz = rz;
            }
        }
        else {
            err = sqrt(rr / st->bb);
            if (err < st->abs_tol || err / st->err0 < st->rel_tol) {
                done = 1;
                converged = 1;
            }
            else if (k - 1 >= st->max_iter) {
                done = 1;
            }
            else {
                beta = rz / sum_partials(pa_prev2, G, 2, 1, lds);
            }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->it = it;
            st->err = err;
            st->rr = rr;
            st->beta = beta;
            if (done) {
                st->done = 1;
                st->converged = converged;
            }
        }
        if (done) return;
        const TV tb = (TV)beta;
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, G, r0, r1);
        double dot = 0.0;
        for (int64_t row = r0 + threadIdx.x; row < r1; row += NT) {
            const int64_t b0 = slice_ptr[row >> 6];
            const int wd = (int)((slice_ptr[(row >> 6) + 1] - b0) >> 6);
            const int lane = (int)(row & 63);
            const int32_t* cp = col + b0 + lane;
            const TS* vp = val + b0 * 9 + lane;
            TV a0 = (TV)0, a1 = (TV)0, a2 = (TV)0;
            for (int j = 0; j < wd; j++) {
                const int64_t c = 3 * (int64_t)cp[j * 64];
                const TV x0 = z[c + 0] + tb * p_old[c + 0];
                const TV x1 = z[c + 1] + tb * p_old[c + 1];
                const TV x2 = z[c + 2] + tb * p_old[c + 2];
                const TS* v = vp + (int64_t)j * 576;
                a0 += (TV)v[0] * x0 + (TV)v[64] * x1 + (TV)v[128] * x2;
                a1 += (TV)v[192] * x0 + (TV)v[256] * x1 + (TV)v[320] * x2;
                a2 += (TV)v[384] * x0 + (TV)v[448] * x1 + (TV)v[512] * x2;
            }
            const int64_t kk = 3 * row;
            const TV p0 = z[kk + 0] + tb * p_old[kk + 0];
            const TV p1 = z[kk + 1] + tb * p_old[kk + 1];
            const TV p2 = z[kk + 2] + tb * p_old[kk + 2];
            p_new[kk + 0] = p0;
            p_new[kk + 1] = p1;
            p_new[kk + 2] = p2;
            q[kk + 0] = a0;
            q[kk + 1] = a1;
            q[kk + 2] = a2;
            dot += (double)p0 * (double)a0 + (double)p1 * (double)a1 + (double)p2 * (double)a2;
        }
        const double s = block_sum(dot, lds);
        if (threadIdx.x == 0) pb[w] = s;
    }

    template <typename TV>
    __global__ __launch_bounds__(NT) void k_pcg_update(int32_t n_rows, int32_t k, const TV* __restrict__ dinv, TV* __restrict__ x,
                                                       TV* __restrict__ r, TV* __restrict__ z, const TV* __restrict__ p,
                                                       const TV* __restrict__ q, PcgState* st, const double* __restrict__ pa_cur,
                                                       double* __restrict__ pa_next, const double* __restrict__ pb)
    {
        __shared__ double lds[NT / 64];
        if (st->done) return;
        const int64_t G = gridDim.x;
        const double pAp = sum_partials(pb, G, 1, 0, lds);
        const double rz = sum_partials(pa_cur, G, 2, 1, lds);
        const double alpha = rz / pAp;
        if (pAp <= 0.0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                st->indef = 1;
                if (st->stop_on_indef) {
                    st->done = 1;
                    st->converged = 0;
                    st->it = k;
                }
            }
            if (st->stop_on_indef) return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->pAp = pAp;
            st->alpha = alpha;
        }
        const TV ta = (TV)alpha;
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, G, r0, r1);
        double rr = 0.0, rzn = 0.0;
        for (int64_t row = r0 + threadIdx.x; row < r1; row += NT) {
            const int64_t kk = 3 * row;
            TV rv[3], zv[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                x[kk + i] = x[kk + i] + ta * p[kk + i];
                rv[i] = r[kk + i] - ta * q[kk + i];
                r[kk + i] = rv[i];
            }
            apply_dinv(dinv, n_rows, row, rv, zv);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                z[kk + i] = zv[i];
                rr += (double)rv[i] * (double)rv[i];
                rzn += (double)rv[i] * (double)zv[i];
            }
        }
        const double srr = block_sum(rr, lds);
        const double srz = block_sum(rzn, lds);
        if (threadIdx.x == 0) {
            pa_next[2 * w + 0] = srr;
            pa_next[2 * w + 1] = srz;
        }
    }

    int pcg_grid(int32_t n_rows) { return reduction_groups(n_rows); }

    void launch_pcg_init64(hipStream_t s, const SellPattern& P, const double* dinv, const double* b, double* x, double* r, double* z,
                           double* p0, double* partials)
    {
        k_pcg_init<double><<<pcg_grid(P.n_rows), NT, 0, s>>>(P.n_rows, dinv, b, x, r, z, p0, partials);
    }
    void launch_pcg_iter64(hipStream_t s, const SellPattern& P, int32_t k, const double* val, const double* dinv, double* x, double* r,
                           double* z, const double* p_old, double* p_new, double* q, PcgState* st, double* partials)
    {
        const int G = pcg_grid(P.n_rows);
        double* pa[2] = { partials, partials + 2 * G };
        double* pb = partials + 4 * G;
        k_pcg_spmv<double, double><<<G, NT, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_col, val, z, p_old, p_new, q, st, pa[(k - 1) & 1],
                                                    pa[k & 1], pb);
        k_pcg_update<double><<<G, NT, 0, s>>>(P.n_rows, k, dinv, x, r, z, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
    }
    void launch_pcg_init32(hipStream_t s, const SellPattern& P, const float* dinv, const float* b, float* x, float* r, float* z,
                           float* p0, double* partials)
    {
        k_pcg_init<float><<<pcg_grid(P.n_rows), NT, 0, s>>>(P.n_rows, dinv, b, x, r, z, p0, partials);
    }
    void launch_pcg_iter32(hipStream_t s, const SellPattern& P, int32_t k, const float* val, const float* dinv, float* x, float* r,
                           float* z, const float* p_old, float* p_new, float* q, PcgState* st, double* partials)
    {
        const int G = pcg_grid(P.n_rows);
        double* pa[2] = { partials, partials + 2 * G };
        double* pb = partials + 4 * G;
        k_pcg_spmv<float, float><<<G, NT, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_col, val, z, p_old, p_new, q, st, pa[(k - 1) & 1],
                                                  pa[k & 1], pb);
        k_pcg_update<float><<<G, NT, 0, s>>>(P.n_rows, k, dinv, x, r, z, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
    }

    // ---- mixed-precision refinement -----------------------------------------------
    // r = b - A x with FP32 matrix values, FP64 vectors and FP64 accumulation (the
    // outer residual of iterative refinement); r32 = (float) r; *rr_out = ||r||^2.
    __global__ __launch_bounds__(NT) void k_residual_mixed(int32_t n_rows, const int64_t* __restrict__ slice_ptr,
                                                           const int32_t* __restrict__ col, const float* __restrict__ val,
                                                           const double* __restrict__ b, const double* __restrict__ x,
                                                           double* __restrict__ r, float* __restrict__ r32, double* __restrict__ partials)
    {
        __shared__ double lds[NT / 64];
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, gridDim.x, r0, r1);
        double rr = 0.0;
        for (int64_t row = r0 + threadIdx.x; row < r1; row += NT) {
            const int64_t b0 = slice_ptr[row >> 6];
            const int wd = (int)((slice_ptr[(row >> 6) + 1] - b0) >> 6);
            const int lane = (int)(row & 63);
            const int32_t* cp = col + b0 + lane;
            const float* vp = val + b0 * 9 + lane;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int j = 0; j < wd; j++) {
                const int64_t c = 3 * (int64_t)cp[j * 64];
                const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
                const float* v = vp + (int64_t)j * 576;
                a0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2;
                a1 += (double)v[192] * x0 + (double)v[256] * x1 + (double)v[320] * x2;
                a2 += (double)v[384] * x0 + (double)v[448] * x1 + (double)v[512] * x2;
            }
            const int64_t k = 3 * row;
            const double e0 = b[k] - a0, e1 = b[k + 1] - a1, e2 = b[k + 2] - a2;
            r[k] = e0;
            r[k + 1] = e1;
            r[k + 2] = e2;
            r32[k] = (float)e0;
            r32[k + 1] = (float)e1;
            r32[k + 2] = (float)e2;
            rr += e0 * e0 + e1 * e1 + e2 * e2;
        }
        const double s = block_sum(rr, lds);
        if (threadIdx.x == 0) partials[w] = s;
    }
    void launch_residual_mixed(hipStream_t s, const SellPattern& P, const float* val, const double* b, const double* x, double* r,
                               float* r32, double* rr_out, double* partials)
    {
        const int G = pcg_grid(P.n_rows);
        k_residual_mixed<<<G, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_col, val, b, x, r, r32, partials);
        k_sum<<<1, NT, 0, s>>>(partials, G, 1, 0, rr_out);
    }

    __global__ __launch_bounds__(NT) void k_x_update_mixed(int64_t n, double* __restrict__ x, const float* __restrict__ d)
    {
        const int64_t i = xcd_remap() * NT + threadIdx.x;
        if (i < n) x[i] += (double)d[i];
    }
    void launch_x_update_mixed(hipStream_t s, int64_t n, double* x, const float* d)
    {
        if (n > 0) k_x_update_mixed<<<grid_for(n), NT, 0, s>>>(n, x, d);
    }

    // ---- vector operations ---------------------------------------------------------
    __global__ __launch_bounds__(NT) void k_axpy(int64_t n, double* __restrict__ dst, const double* __restrict__ a, double sc,
                                                 const double* __restrict__ b)
    {
        const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (i < n) dst[i] = a[i] + sc * b[i];
    }
    void launch_axpy(hipStream_t s, int64_t n, double* dst, const double* a, double sc, const double* b)
    {
        if (n > 0) k_axpy<<<grid_for(n), NT, 0, s>>>(n, dst, a, sc, b);
    }
    __global__ __launch_bounds__(NT) void k_scale(int64_t n, double* __restrict__ dst, const double* __restrict__ a, double sc)
    {
        const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (i < n) dst[i] = sc * a[i];
    }
    void launch_scale(hipStream_t s, int64_t n, double* dst, const double* a, double sc)
    {
        if (n > 0) k_scale<<<grid_for(n), NT, 0, s>>>(n, dst, a, sc);
    }

    __global__ __launch_bounds__(NT) void k_dot_partials(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                         double* __restrict__ partials)
    {
        __shared__ double lds[NT / 64];
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n, w, gridDim.x, r0, r1);
        double acc = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) acc += a[i] * b[i];
        const double s = block_sum(acc, lds);
        if (threadIdx.x == 0) partials[w] = s;
    }
    void launch_dot(hipStream_t s, int64_t n, const double* a, const double* b, double* out, double* partials)
    {
        const int G = reduction_groups(n);
        k_dot_partials<<<G, NT, 0, s>>>(n, a, b, partials);
        k_sum<<<1, NT, 0, s>>>(partials, G, 1, 0, out);
    }

    __global__ __launch_bounds__(NT) void k_max_abs_partials(int64_t n, const double* __restrict__ a, double* __restrict__ partials)
    {
        __shared__ double lds[NT / 64];
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n, w, gridDim.x, r0, r1);
        double m = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) m = nanmax(m, fabs(a[i]));
        const double s = block_max(m, lds);
        if (threadIdx.x == 0) partials[w] = s;
    }
    __global__ __launch_bounds__(NT) void k_max_final(const double* __restrict__ partials, int64_t G, double* out)
    {
        __shared__ double lds[NT / 64];
        double m = 0.0;
        for (int64_t i = threadIdx.x; i < G; i += NT) m = nanmax(m, partials[i]);
        const double r = block_max(m, lds);
        if (threadIdx.x == 0) *out = r;
    }
    void launch_max_abs(hipStream_t s, int64_t n, const double* a, double* out, double* partials)
    {
        const int G = reduction_groups(n);
        k_max_abs_partials<<<G, NT, 0, s>>>(n, a, partials);
        k_max_final<<<1, NT, 0, s>>>(partials, G, out);
    }

    __global__ __launch_bounds__(NT) void k_fill_u8(int64_t n, uint8_t* p, uint8_t v)
    {
        const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (i < n) p[i] = v;
    }
    void launch_fill_u8(hipStream_t s, int64_t n, uint8_t* p, uint8_t v)
    {
        if (n > 0) k_fill_u8<<<grid_for(n), NT, 0, s>>>(n, p, v);
    }
}
