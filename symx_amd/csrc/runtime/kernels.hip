// Ahead-of-time kernels of the backend (gfx950): reductions, PD projection,
// deterministic assembly into the sliced block matrix, block-Jacobi PCG and the
// Newton vector operations. Generated element kernels live in hiprtc modules.
//
// Every reduction is two-level with a fixed order: each workgroup writes one
// partial, and the last workgroup to finish (found with an integer ticket)
// sums the partials in index order. No floating-point atomics are used, so
// results are bitwise repeatable for a given problem size.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.h"

namespace symx_hip
{
    constexpr int NT = 256;           // threads per workgroup
    constexpr int MAX_RED_GROUPS = 1024;

    int reduction_groups(int64_t n)
    {
        int64_t g = (n + NT - 1) / NT;
        if (g < 1) g = 1;
        if (g > MAX_RED_GROUPS) g = MAX_RED_GROUPS;
        return (int)g;
    }

    static inline unsigned grid_for(int64_t n) { return (unsigned)((n + NT - 1) / NT > 0 ? (n + NT - 1) / NT : 1); }

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

    // Sum over the workgroup; the result is valid in thread 0.
    __device__ inline double block_sum(double v, double* lds)
    {
        v = wave_sum(v);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) lds[w] = v;
        __syncthreads();
        double r = 0.0;
        if (threadIdx.x == 0) {
            for (int i = 0; i < NT / 64; i++) r += lds[i];
        }
        __syncthreads();
        return r;
    }
    __device__ inline double block_max(double v, double* lds)
    {
        v = wave_max(v);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) lds[w] = v;
        __syncthreads();
        double r = 0.0;
        if (threadIdx.x == 0) {
            r = lds[0];
            for (int i = 1; i < NT / 64; i++) r = nanmax(r, lds[i]);
        }
        __syncthreads();
        return r;
    }

    __device__ inline void store_coherent(double* p, double v)
    {
        __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ inline double load_coherent(const double* p)
    {
        return __longlong_as_double((long long)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT));
    }

    // Publish this workgroup's partials (already stored by thread 0) and find out
    // whether it is the last workgroup of the grid. Resets the ticket when it is.
    __device__ inline bool last_group(unsigned int* ticket)
    {
        __shared__ int is_last;
        __threadfence();
        if (threadIdx.x == 0) {
            const unsigned t = atomicAdd(ticket, 1u);
            is_last = (t == gridDim.x - 1) ? 1 : 0;
            if (is_last) *ticket = 0u;
        }
        __syncthreads();
        if (is_last) __threadfence();
        return is_last != 0;
    }

    // Fixed-order sum of `n` partials by one workgroup (thread 0 gets the result).
    __device__ inline double sum_partials(const double* partials, int n, int stride, int offset, double* lds)
    {
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += NT) acc += load_coherent(partials + (int64_t)i * stride + offset);
        return block_sum(acc, lds);
    }

    // ---- plain sum ---------------------------------------------------------------
    __global__ __launch_bounds__(NT) void k_sum(const double* __restrict__ in, int64_t n, double* out)
    {
        __shared__ double lds[NT / 64];
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += NT) acc += in[i];
        const double r = block_sum(acc, lds);
        if (threadIdx.x == 0) *out = r;
    }
    void launch_sum(hipStream_t s, const double* d_in, int64_t n, double* d_out)
    {
        k_sum<<<1, NT, 0, s>>>(d_in, n, d_out);
    }

    // ---- gradient assembly: grad[3b+i] = sum of element gradient blocks -----------
    __global__ __launch_bounds__(NT) void k_grad_assemble(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                          const double* __restrict__ G, double* __restrict__ grad, int32_t nb)
    {
        const int b = blockIdx.x * NT + threadIdx.x;
        if (b >= nb) return;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        const uint32_t k1 = off[b + 1];
        for (uint32_t k = off[b]; k < k1; k++) {
            const double* src = G + (int64_t)list[k] * 3;
            g0 += src[0];
            g1 += src[1];
            g2 += src[2];
        }
        grad[3 * (int64_t)b + 0] = g0;
        grad[3 * (int64_t)b + 1] = g1;
        grad[3 * (int64_t)b + 2] = g2;
    }
    void launch_grad_assemble(hipStream_t s, const uint32_t* off, const uint32_t* list, const double* G, double* grad, int32_t nb)
    {
        if (nb <= 0) return;
        k_grad_assemble<<<grid_for(nb), NT, 0, s>>>(off, list, G, grad, nb);
    }

    // ---- PD projection -------------------------------------------------------------
    // Cyclic Jacobi eigen-decomposition of a symmetric N x N matrix held in registers
    // (upper triangle only), followed by A <- V f(Lambda) V^T with f(l) = eps (clamp)
    // or -l (mirror) for l < eps, like project_to_PD.cpp:13-32. Returns true if the
    // matrix was rebuilt.
    template <int N>
    struct Sym
    {
        static constexpr int size = N * (N + 1) / 2;
        __device__ static constexpr int idx(int i, int j)
        {
            return (i <= j) ? (i * N - (i * (i - 1)) / 2 + (j - i)) : (j * N - (j * (j - 1)) / 2 + (i - j));
        }
    };

    template <int N>
    __device__ inline bool jacobi_project(double (&a)[Sym<N>::size], double eps, bool mirror)
    {
        double v[N][N];
#pragma unroll
        for (int i = 0; i < N; i++) {
#pragma unroll
            for (int j = 0; j < N; j++) v[i][j] = (i == j) ? 1.0 : 0.0;
        }
        double fro = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
#pragma unroll
            for (int j = i; j < N; j++) fro += (i == j ? 1.0 : 2.0) * a[Sym<N>::idx(i, j)] * a[Sym<N>::idx(i, j)];
        }
        const double tol = 1e-30 * fro;
        for (int sweep = 0; sweep < 24; sweep++) {
            double off = 0.0;
#pragma unroll
            for (int p = 0; p < N; p++) {
#pragma unroll
                for (int q = p + 1; q < N; q++) off += a[Sym<N>::idx(p, q)] * a[Sym<N>::idx(p, q)];
            }
            if (off <= tol) break;
#pragma unroll
            for (int p = 0; p < N; p++) {
#pragma unroll
                for (int q = p + 1; q < N; q++) {
                    const double apq = a[Sym<N>::idx(p, q)];
                    if (fabs(apq) > 1e-300) {
                        const double app = a[Sym<N>::idx(p, p)];
                        const double aqq = a[Sym<N>::idx(q, q)];
                        const double theta = (aqq - app) / (2.0 * apq);
                        double t;
                        if (fabs(theta) > 1e150) {
                            t = 0.5 / theta;
                        }
                        else {
                            t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                            if (theta < 0.0) t = -t;
                        }
                        const double c = 1.0 / sqrt(t * t + 1.0);
                        const double s = t * c;
                        const double tau = s / (1.0 + c);
                        a[Sym<N>::idx(p, p)] = app - t * apq;
                        a[Sym<N>::idx(q, q)] = aqq + t * apq;
                        a[Sym<N>::idx(p, q)] = 0.0;
#pragma unroll
                        for (int r = 0; r < N; r++) {
                            if (r != p && r != q) {
                                const double g = a[Sym<N>::idx(r, p)];
                                const double h = a[Sym<N>::idx(r, q)];
                                a[Sym<N>::idx(r, p)] = g - s * (h + g * tau);
                                a[Sym<N>::idx(r, q)] = h + s * (g - h * tau);
                            }
                        }
#pragma unroll
                        for (int r = 0; r < N; r++) {
                            const double g = v[r][p];
                            const double h = v[r][q];
                            v[r][p] = g - s * (h + g * tau);
                            v[r][q] = h + s * (g - h * tau);
                        }
                    }
                }
            }
        }
        double lam[N];
        bool changed = false;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double l = a[Sym<N>::idx(i, i)];
            if (l < eps) {
                changed = true;
                lam[i] = mirror ? -l : eps;
            }
            else {
                lam[i] = l;
            }
        }
        if (changed) {
#pragma unroll
            for (int i = 0; i < N; i++) {
#pragma unroll
                for (int j = i; j < N; j++) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < N; k++) acc += v[i][k] * lam[k] * v[j][k];
                    a[Sym<N>::idx(i, j)] = acc;
                }
            }
        }
        return changed;
    }

    // Cholesky of A + delta I (upper triangle, copy). True if every pivot is positive,
    // i.e. the smallest eigenvalue of A is above about -delta.
    template <int N>
    __device__ inline bool shifted_cholesky_ok(const double (&a)[Sym<N>::size], double delta)
    {
        double l[Sym<N>::size];
#pragma unroll
        for (int i = 0; i < Sym<N>::size; i++) l[i] = a[i];
#pragma unroll
        for (int i = 0; i < N; i++) l[Sym<N>::idx(i, i)] += delta;
#pragma unroll
        for (int j = 0; j < N; j++) {
            double d = l[Sym<N>::idx(j, j)];
#pragma unroll
            for (int k = 0; k < j; k++) d -= l[Sym<N>::idx(k, j)] * l[Sym<N>::idx(k, j)];
            if (!(d > 0.0)) return false;
            const double rd = 1.0 / sqrt(d);
            l[Sym<N>::idx(j, j)] = d * rd;
#pragma unroll
            for (int i = j + 1; i < N; i++) {
                double x = l[Sym<N>::idx(j, i)];
#pragma unroll
                for (int k = 0; k < j; k++) x -= l[Sym<N>::idx(k, j)] * l[Sym<N>::idx(k, i)];
                l[Sym<N>::idx(j, i)] = x * rd;
            }
        }
        return true;
    }

    // Element Hessians stored as upper 3x3 blocks (a <= b), 9 row-major values each.
    template <int NB>
    __global__ __launch_bounds__(64) void k_project_ub(double* __restrict__ H, int64_t n_mats, double eps, int mirror,
                                                      double skip_tol, unsigned long long* count)
    {
        constexpr int N = 3 * NB;
        constexpr int NUB = NB * (NB + 1) / 2;
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        bool changed = false;
        if (e < n_mats) {
            double* h = H + e * (NUB * 9);
            double a[Sym<N>::size];
            int ub = 0;
#pragma unroll
            for (int bi = 0; bi < NB; bi++) {
#pragma unroll
                for (int bj = bi; bj < NB; bj++) {
#pragma unroll
                    for (int i = 0; i < 3; i++) {
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            if (bi != bj || i <= j) a[Sym<N>::idx(3 * bi + i, 3 * bj + j)] = h[ub * 9 + 3 * i + j];
                        }
                    }
                    ub++;
                }
            }
            bool skip = false;
            if (skip_tol > 0.0) {
                double dmax = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) dmax = fmax(dmax, fabs(a[Sym<N>::idx(i, i)]));
                skip = shifted_cholesky_ok<N>(a, skip_tol * dmax + eps);
            }
            if (!skip) {
                changed = jacobi_project<N>(a, eps, mirror != 0);
                if (changed) {
                    ub = 0;
#pragma unroll
                    for (int bi = 0; bi < NB; bi++) {
#pragma unroll
                        for (int bj = bi; bj < NB; bj++) {
#pragma unroll
                            for (int i = 0; i < 3; i++) {
#pragma unroll
                                for (int j = 0; j < 3; j++) h[ub * 9 + 3 * i + j] = a[Sym<N>::idx(3 * bi + i, 3 * bj + j)];
                            }
                            ub++;
                        }
                    }
                }
            }
        }
        if (count != nullptr) {
            const unsigned long long m = __ballot(changed);
            if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(count, (unsigned long long)__popcll(m));
        }
    }

    int launch_project_ub(hipStream_t s, double* H, int64_t n_mats, int32_t nb, double eps, int32_t mirror, double skip_tol,
                          unsigned long long* d_count)
    {
        if (n_mats <= 0) return 0;
        const unsigned g = (unsigned)((n_mats + 63) / 64);
        switch (nb) {
        case 1: k_project_ub<1><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, skip_tol, d_count); return 0;
        case 2: k_project_ub<2><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, skip_tol, d_count); return 0;
        case 3: k_project_ub<3><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, skip_tol, d_count); return 0;
        case 4: k_project_ub<4><<<g, 64, 0, s>>>(H, n_mats, eps, mirror, skip_tol, d_count); return 0;
        default: return -1;
        }
    }

    template <int N>
    __global__ __launch_bounds__(64) void k_project_dense(double* __restrict__ M, int64_t n_mats, double eps, int mirror,
                                                         uint8_t* changed_out)
    {
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        if (e >= n_mats) return;
        double* m = M + e * (N * N);
        double a[Sym<N>::size];
#pragma unroll
        for (int i = 0; i < N; i++) {
#pragma unroll
            for (int j = i; j < N; j++) a[Sym<N>::idx(i, j)] = m[i * N + j];
        }
        const bool changed = jacobi_project<N>(a, eps, mirror != 0);
        if (changed) {
#pragma unroll
            for (int i = 0; i < N; i++) {
#pragma unroll
                for (int j = 0; j < N; j++) m[i * N + j] = a[Sym<N>::idx(i, j)];
            }
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
        const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
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
        const int row = blockIdx.x * NT + threadIdx.x;
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
    // Iteration = two kernels:
    //   k_pcg_spmv:   p_new = z + beta p_old (own row, and on the fly for neighbours),
    //                 q = A p_new, partial p.q; last group: alpha = rz / pAp.
    //   k_pcg_update: x += alpha p, r -= alpha q, z = D^-1 r, partial r.r and r.z;
    //                 last group: convergence test of solve_pcg.h:200-210, beta.
    // Both return at once when the solve is done, so the host can enqueue batches.

    template <typename T>
    __global__ __launch_bounds__(NT) void k_pcg_init(int32_t n_rows, const T* __restrict__ dinv, const T* __restrict__ b,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, T* __restrict__ p0,
                                                     PcgState* st, double* partials, unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        const int row = blockIdx.x * NT + threadIdx.x;
        double bb = 0.0, rz = 0.0;
        if (row < n_rows) {
            T rv[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int64_t k = 3 * (int64_t)row + i;
                rv[i] = b[k];
                x[k] = (T)0;
                r[k] = rv[i];
                p0[k] = (T)0;
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                T zi = dinv[(int64_t)(3 * i + 0) * n_rows + row] * rv[0] + dinv[(int64_t)(3 * i + 1) * n_rows + row] * rv[1] +
                       dinv[(int64_t)(3 * i + 2) * n_rows + row] * rv[2];
                z[3 * (int64_t)row + i] = zi;
                bb += (double)rv[i] * (double)rv[i];
                rz += (double)rv[i] * (double)zi;
            }
        }
        const double sbb = block_sum(bb, lds);
        const double srz = block_sum(rz, lds);
        if (threadIdx.x == 0) {
            store_coherent(partials + 2 * blockIdx.x + 0, sbb);
            store_coherent(partials + 2 * blockIdx.x + 1, srz);
        }
        if (!last_group(ticket)) return;
        const double tbb = sum_partials(partials, gridDim.x, 2, 0, lds);
        const double trz = sum_partials(partials, gridDim.x, 2, 1, lds);
        if (threadIdx.x == 0) {
            st->bb = tbb;
            st->rz = trz;
            st->beta = 0.0;
            st->it = 0;
            st->indef = 0;
            st->done = 0;
            st->converged = 0;
            if (tbb < st->abs_tol * st->abs_tol) {
                st->done = 1;
                st->converged = 1;
                st->err = 0.0;
                st->err0 = 0.0;
            }
            else {
                // x = 0 so r = b and ||r|| / ||b|| is exactly sqrt(bb / bb).
                const double err = sqrt(tbb / tbb);
                st->rr = tbb;
                st->err = err;
                st->err0 = err;
                if (err < st->abs_tol) {
                    st->done = 1;
                    st->converged = 1;
                }
                else if (st->max_iter <= 0) {
                    st->done = 1;
                }
            }
        }
    }

    template <typename TS, typename TV>
    __global__ __launch_bounds__(NT) void k_pcg_spmv(int32_t n_rows, const int64_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ col, const TS* __restrict__ val,
                                                     const TV* __restrict__ z, const TV* __restrict__ p_old, TV* __restrict__ p_new,
                                                     TV* __restrict__ q, PcgState* st, double* partials, unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        if (st->done) return;
        const TV beta = (TV)st->beta;
        const int row = blockIdx.x * NT + threadIdx.x;
        double dot = 0.0;
        if (row < n_rows) {
            const int64_t b0 = slice_ptr[row >> 6];
            const int w = (int)((slice_ptr[(row >> 6) + 1] - b0) >> 6);
            const int lane = row & 63;
            const int32_t* cp = col + b0 + lane;
            const TS* vp = val + b0 * 9 + lane;
            TV a0 = (TV)0, a1 = (TV)0, a2 = (TV)0;
            for (int j = 0; j < w; j++) {
                const int64_t c = 3 * (int64_t)cp[j * 64];
                const TV x0 = z[c + 0] + beta * p_old[c + 0];
                const TV x1 = z[c + 1] + beta * p_old[c + 1];
                const TV x2 = z[c + 2] + beta * p_old[c + 2];
                const TS* v = vp + (int64_t)j * 576;
                a0 += (TV)v[0] * x0 + (TV)v[64] * x1 + (TV)v[128] * x2;
                a1 += (TV)v[192] * x0 + (TV)v[256] * x1 + (TV)v[320] * x2;
                a2 += (TV)v[384] * x0 + (TV)v[448] * x1 + (TV)v[512] * x2;
            }
            const int64_t k = 3 * (int64_t)row;
            const TV p0 = z[k + 0] + beta * p_old[k + 0];
            const TV p1 = z[k + 1] + beta * p_old[k + 1];
            const TV p2 = z[k + 2] + beta * p_old[k + 2];
            p_new[k + 0] = p0;
            p_new[k + 1] = p1;
            p_new[k + 2] = p2;
            q[k + 0] = a0;
            q[k + 1] = a1;
            q[k + 2] = a2;
            dot = (double)p0 * (double)a0 + (double)p1 * (double)a1 + (double)p2 * (double)a2;
        }
        const double s = block_sum(dot, lds);
        if (threadIdx.x == 0) store_coherent(partials + blockIdx.x, s);
        if (!last_group(ticket)) return;
        const double pAp = sum_partials(partials, gridDim.x, 1, 0, lds);
        if (threadIdx.x == 0) {
            st->pAp = pAp;
            st->alpha = st->rz / pAp;
            if (pAp <= 0.0) {
                st->indef = 1;
                if (st->stop_on_indef) {
                    st->done = 1;
                    st->converged = 0;
                    st->it = st->it + 1;
                }
            }
        }
    }

    template <typename TV>
    __global__ __launch_bounds__(NT) void k_pcg_update(int32_t n_rows, const TV* __restrict__ dinv, TV* __restrict__ x,
                                                       TV* __restrict__ r, TV* __restrict__ z, const TV* __restrict__ p,
                                                       const TV* __restrict__ q, PcgState* st, double* partials, unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        if (st->done) return;
        const TV alpha = (TV)st->alpha;
        const int row = blockIdx.x * NT + threadIdx.x;
        double rr = 0.0, rz = 0.0;
        if (row < n_rows) {
            const int64_t k = 3 * (int64_t)row;
            TV rv[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                x[k + i] = x[k + i] + alpha * p[k + i];
                rv[i] = r[k + i] - alpha * q[k + i];
                r[k + i] = rv[i];
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const TV zi = dinv[(int64_t)(3 * i + 0) * n_rows + row] * rv[0] + dinv[(int64_t)(3 * i + 1) * n_rows + row] * rv[1] +
                              dinv[(int64_t)(3 * i + 2) * n_rows + row] * rv[2];
                z[k + i] = zi;
                rr += (double)rv[i] * (double)rv[i];
                rz += (double)rv[i] * (double)zi;
            }
        }
        const double srr = block_sum(rr, lds);
        const double srz = block_sum(rz, lds);
        if (threadIdx.x == 0) {
            store_coherent(partials + 2 * blockIdx.x + 0, srr);
            store_coherent(partials + 2 * blockIdx.x + 1, srz);
        }
        if (!last_group(ticket)) return;
        const double trr = sum_partials(partials, gridDim.x, 2, 0, lds);
        const double trz = sum_partials(partials, gridDim.x, 2, 1, lds);
        if (threadIdx.x == 0) {
            const int it = st->it + 1;
            const double err = sqrt(trr / st->bb);
            st->rr = trr;
            st->err = err;
            st->it = it;
            if (err < st->abs_tol || err / st->err0 < st->rel_tol) {
                st->done = 1;
                st->converged = 1;
            }
            else if (it >= st->max_iter) {
                st->done = 1;
                st->converged = 0;
            }
            else {
                st->beta = trz / st->rz;
                st->rz = trz;
            }
        }
    }

    void launch_pcg_init64(hipStream_t s, const SellPattern& P, const double* dinv, const double* b, double* x, double* r, double* z,
                           double* p0, PcgState* st, double* partials, unsigned int* ticket)
    {
        k_pcg_init<double><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, dinv, b, x, r, z, p0, st, partials, ticket);
    }
    void launch_pcg_iter64(hipStream_t s, const SellPattern& P, const double* val, const double* dinv, double* x, double* r, double* z,
                           const double* p_old, double* p_new, double* q, PcgState* st, double* partials, unsigned int* ticket)
    {
        const unsigned g = grid_for(P.n_rows);
        k_pcg_spmv<double, double><<<g, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_col, val, z, p_old, p_new, q, st, partials, ticket);
        k_pcg_update<double><<<g, NT, 0, s>>>(P.n_rows, dinv, x, r, z, p_new, q, st, partials, ticket);
    }
    void launch_pcg_init32(hipStream_t s, const SellPattern& P, const float* dinv, const float* b, float* x, float* r, float* z,
                           float* p0, PcgState* st, double* partials, unsigned int* ticket)
    {
        k_pcg_init<float><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, dinv, b, x, r, z, p0, st, partials, ticket);
    }
    void launch_pcg_iter32(hipStream_t s, const SellPattern& P, const float* val, const float* dinv, float* x, float* r, float* z,
                           const float* p_old, float* p_new, float* q, PcgState* st, double* partials, unsigned int* ticket)
    {
        const unsigned g = grid_for(P.n_rows);
        k_pcg_spmv<float, float><<<g, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_col, val, z, p_old, p_new, q, st, partials, ticket);
        k_pcg_update<float><<<g, NT, 0, s>>>(P.n_rows, dinv, x, r, z, p_new, q, st, partials, ticket);
    }

    // ---- mixed-precision refinement -----------------------------------------------
    // r = b - A x with FP32 matrix values, FP64 vectors and FP64 accumulation (the
    // outer residual of iterative refinement); r32 = (float) r; *rr_out = ||r||^2.
    __global__ __launch_bounds__(NT) void k_residual_mixed(int32_t n_rows, const int64_t* __restrict__ slice_ptr,
                                                           const int32_t* __restrict__ col, const float* __restrict__ val,
                                                           const double* __restrict__ b, const double* __restrict__ x,
                                                           double* __restrict__ r, float* __restrict__ r32, double* rr_out,
                                                           double* partials, unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        const int row = blockIdx.x * NT + threadIdx.x;
        double rr = 0.0;
        if (row < n_rows) {
            const int64_t b0 = slice_ptr[row >> 6];
            const int w = (int)((slice_ptr[(row >> 6) + 1] - b0) >> 6);
            const int lane = row & 63;
            const int32_t* cp = col + b0 + lane;
            const float* vp = val + b0 * 9 + lane;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int j = 0; j < w; j++) {
                const int64_t c = 3 * (int64_t)cp[j * 64];
                const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
                const float* v = vp + (int64_t)j * 576;
                a0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2;
                a1 += (double)v[192] * x0 + (double)v[256] * x1 + (double)v[320] * x2;
                a2 += (double)v[384] * x0 + (double)v[448] * x1 + (double)v[512] * x2;
            }
            const int64_t k = 3 * (int64_t)row;
            const double r0 = b[k] - a0, r1 = b[k + 1] - a1, r2 = b[k + 2] - a2;
            r[k] = r0;
            r[k + 1] = r1;
            r[k + 2] = r2;
            r32[k] = (float)r0;
            r32[k + 1] = (float)r1;
            r32[k + 2] = (float)r2;
            rr = r0 * r0 + r1 * r1 + r2 * r2;
        }
        const double s = block_sum(rr, lds);
        if (threadIdx.x == 0) store_coherent(partials + blockIdx.x, s);
        if (!last_group(ticket)) return;
        const double t = sum_partials(partials, gridDim.x, 1, 0, lds);
        if (threadIdx.x == 0) *rr_out = t;
    }
    void launch_residual_mixed(hipStream_t s, const SellPattern& P, const float* val, const double* b, const double* x, double* r,
                               float* r32, double* rr_out, double* partials, unsigned int* ticket)
    {
        k_residual_mixed<<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_col, val, b, x, r, r32, rr_out, partials,
                                                          ticket);
    }

    __global__ __launch_bounds__(NT) void k_x_update_mixed(int64_t n, double* __restrict__ x, const float* __restrict__ d)
    {
        const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
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

    __global__ __launch_bounds__(NT) void k_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* out,
                                                double* partials, unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        double acc = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) acc += a[i] * b[i];
        const double s = block_sum(acc, lds);
        if (threadIdx.x == 0) store_coherent(partials + blockIdx.x, s);
        if (!last_group(ticket)) return;
        const double t = sum_partials(partials, gridDim.x, 1, 0, lds);
        if (threadIdx.x == 0) *out = t;
    }
    void launch_dot(hipStream_t s, int64_t n, const double* a, const double* b, double* out, double* partials, unsigned int* ticket)
    {
        k_dot<<<reduction_groups(n), NT, 0, s>>>(n, a, b, out, partials, ticket);
    }

    __global__ __launch_bounds__(NT) void k_max_abs(int64_t n, const double* __restrict__ a, double* out, double* partials,
                                                    unsigned int* ticket)
    {
        __shared__ double lds[NT / 64];
        double m = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
            m = nanmax(m, fabs(a[i]));
        }
        const double s = block_max(m, lds);
        if (threadIdx.x == 0) store_coherent(partials + blockIdx.x, s);
        if (!last_group(ticket)) return;
        double acc = 0.0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += NT) {
            acc = nanmax(acc, load_coherent(partials + i));
        }
        const double t = block_max(acc, lds);
        if (threadIdx.x == 0) *out = t;
    }
    void launch_max_abs(hipStream_t s, int64_t n, const double* a, double* out, double* partials, unsigned int* ticket)
    {
        k_max_abs<<<reduction_groups(n), NT, 0, s>>>(n, a, out, partials, ticket);
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
