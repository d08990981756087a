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
#include <mutex>
#include <type_traits>

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

    // Matrix-value load: non-temporal when the matrix is streamed (larger than the
    // last-level caches, so it is read from HBM anyway and must not displace the
    // gathered vector entries), ordinary when it stays cache resident across
    // iterations (small systems, where a non-temporal read would go to HBM each time).
    template <bool STREAM, typename T>
    __device__ __forceinline__ T mload(const T* p)
    {
        if constexpr (STREAM) return __builtin_nontemporal_load(p);
        else return *p;
    }
    // The nine values of the 3x3 block in `lane` of a 64-slot value group g (layout of
    // sell_val_index): runs of SellRun<TS>::width values as one load each, then value 8.
    template <typename TS, bool STREAM>
    __device__ __forceinline__ void load_block(const TS* __restrict__ g, int lane, TS (&v)[9])
    {
        constexpr int W = SellRun<TS>::width;
        if constexpr (W == 1) {
#pragma unroll
            for (int q = 0; q < 9; q++) v[q] = mload<STREAM>(g + 64 * q + lane);
        }
        else {
            typedef TS run_t __attribute__((ext_vector_type(W)));
#pragma unroll
            for (int r = 0; r < 8 / W; r++) {
                const run_t t = mload<STREAM>(reinterpret_cast<const run_t*>(g + r * 64 * W) + lane);
#pragma unroll
                for (int u = 0; u < W; u++) v[r * W + u] = t[u];
            }
            v[8] = mload<STREAM>(g + 512 + lane);
        }
    }
    // Block column of a SELL slot: int32 ids, or int16 offsets from the slot's row (half
    // the index bytes; used when every offset fits).
    template <typename CT>
    __device__ __forceinline__ int64_t block_col(CT stored, int64_t row)
    {
        if constexpr (sizeof(CT) == 2) return row + (int64_t)stored;
        else return (int64_t)stored;
    }
    // SpMV and PCG-update workgroup size: kSpmvThreads (internal.h).
    // Mean blocks per row from which the search direction gets a pass of its own.
    constexpr int kDirectionPassMeanRow = 32;
    // Matrices above this many value bytes are streamed (the MI355X infinity cache is 256 MiB).
    constexpr double kStreamBytes = 128.0 * 1024 * 1024;

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

    // Workgroup sum in a fixed order, returned to every thread (NTS threads, lds[NTS / 64]).
    template <int NTS = NT>
    __device__ inline double block_sum(double v, double* lds)
    {
        v = wave_sum(v);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) lds[w] = v;
        __syncthreads();
        double r = 0.0;
        for (int i = 0; i < NTS / 64; i++) r += lds[i];
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

    // Fixed-order sum of G partials p[i*stride + off], returned to every thread. The first
    // NT threads accumulate and further waves add exact zeros, so a workgroup of any
    // NTS >= NT threads forms the same value bit for bit as an NT-thread one.
    template <int NTS = NT>
    __device__ inline double sum_partials(const double* __restrict__ p, int64_t G, int stride, int off, double* lds)
    {
        double acc = 0.0;
        if (threadIdx.x < NT)
            for (int64_t i = threadIdx.x; i < G; i += NT) acc += p[i * stride + off];
        return block_sum<NTS>(acc, lds);
    }

    // K sums of G <= MAX_GROUPS partials at once, out[c] = sum of p[c][i * stride[c]], each
    // formed in the order of sum_partials (bitwise the same value). All loads are issued
    // before the first add and the K workgroup reductions share one pair of barriers, so
    // the latency is that of one sum (lds: K * NTS / 64 doubles).
    constexpr int kPartialsPerThread = (MAX_GROUPS + NT - 1) / NT;
    template <int NTS, int K>
    __device__ inline void sum_partials_k(const double* const (&p)[K], const int (&stride)[K], int64_t G, double* lds, double (&out)[K])
    {
        double v[K][kPartialsPerThread];
#pragma unroll
        for (int c = 0; c < K; c++)
#pragma unroll
            for (int u = 0; u < kPartialsPerThread; u++) {
                const int64_t i = threadIdx.x + (int64_t)u * NT;
                v[c][u] = (threadIdx.x < NT && i < G) ? p[c][i * stride[c]] : 0.0;
            }
        double acc[K];
#pragma unroll
        for (int c = 0; c < K; c++) {
            acc[c] = 0.0;
#pragma unroll
            for (int u = 0; u < kPartialsPerThread; u++)
                if (threadIdx.x < NT && threadIdx.x + (int64_t)u * NT < G) acc[c] += v[c][u];
            acc[c] = wave_sum(acc[c]);
        }
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int c = 0; c < K; c++) lds[c * (NTS / 64) + w] = acc[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < K; c++) {
            double r = 0.0;
            for (int i = 0; i < NTS / 64; i++) r += lds[c * (NTS / 64) + i];
            out[c] = r;
        }
    }

    // K workgroup sums with one pair of barriers, each bitwise equal to block_sum's
    // (lds: K * NTS / 64 doubles).
    template <int NTS, int K>
    __device__ inline void block_sum_k(double (&v)[K], double* lds)
    {
#pragma unroll
        for (int c = 0; c < K; c++) v[c] = wave_sum(v[c]);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int c = 0; c < K; c++) lds[c * (NTS / 64) + w] = v[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < K; c++) {
            double t = 0.0;
            for (int i = 0; i < NTS / 64; i++) t += lds[c * (NTS / 64) + i];
            v[c] = t;
        }
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
    // List t belongs to DoF block rows[t] (lists stored in the assembly traversal order).
    // One lane per component: a wave sums 21 lists (63 lanes), so each gathered 24-byte
    // block is one contiguous read of three lanes, and a third as many lists are in flight
    // as with a thread per list (their element blocks then stay in L2 until the lists of
    // the neighbouring rows reach them). Every component is summed in list order, as before.
    constexpr int kGradListsPerWave = 21;
    // List entries whose loads are issued together in the assembly gathers (each entry is a
    // dependent pair: list id, then value). Slot lists hold about 7 element blocks on Tet4
    // meshes, gradient lists about 25. 6M-tet block: assembly 1.63 ms loading one entry at a
    // time, 1.13 ms four at a time (1.17 ms six, 1.29 ms eight); gradient 0.68 ms one at a
    // time, 0.30 / 0.27 / 0.26 ms four / eight / twelve.
    constexpr int kSlotAhead = 4, kGradAhead = 12;
    __global__ __launch_bounds__(NT) void k_grad_assemble(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                          const int32_t* __restrict__ rows, const double* __restrict__ G,
                                                          double* __restrict__ grad, int32_t nb)
    {
        const int lane = threadIdx.x & 63, sub = lane / 3, c = lane - 3 * sub;
        if (sub >= kGradListsPerWave) return;
        const int64_t t = ((xcd_remap() * NT + threadIdx.x) >> 6) * kGradListsPerWave + sub;
        if (t >= nb) return;
        double g = 0.0;
        const uint32_t k1 = off[t + 1];
        uint32_t k = off[t];
        // kGradAhead entries per round: their ids, then their values, load independently
        for (; k + kGradAhead <= k1; k += kGradAhead) {
            uint32_t id[kGradAhead];
            double v[kGradAhead];
#pragma unroll
            for (int u = 0; u < kGradAhead; u++) id[u] = list[k + u];
#pragma unroll
            for (int u = 0; u < kGradAhead; u++) v[u] = G[(int64_t)id[u] * 3 + c];
#pragma unroll
            for (int u = 0; u < kGradAhead; u++) g += v[u];
        }
        for (; k < k1; k++) g += G[(int64_t)list[k] * 3 + c];
        grad[3 * (int64_t)rows[t] + c] = g;
    }
    void launch_grad_assemble(hipStream_t s, const uint32_t* off, const uint32_t* list, const int32_t* rows, const double* G, double* grad,
                              int32_t nb)
    {
        if (nb <= 0) return;
        const int64_t waves = ((int64_t)nb + kGradListsPerWave - 1) / kGradListsPerWave;
        k_grad_assemble<<<(unsigned)((waves + NT / 64 - 1) / (NT / 64)), NT, 0, s>>>(off, list, rows, G, grad, nb);
    }

    // ---- PD projection (device code in projection.h) ---------------------------------
    // Full-size projection of an element Hessian stored as upper 3x3 blocks.
    template <int NB, typename T>
    __device__ __forceinline__ bool project_element_full(T* __restrict__ h, double eps, bool mirror)
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

    // 4-node element Hessian with the rigid translations as null vectors: H = U M U^T with
    // U = W (x) I3, so the projection acts on the 9x9 M = U^T H U and maps the null space
    // with f(0). Only M's eigenpairs below eps are computed (Householder tridiagonalisation,
    // Sturm bisection, inverse iteration, Gram-Schmidt, back-transform), and H is updated as
    //     H + sum_j (f(lambda_j) - lambda_j) z_j z_j^T + f(0) T T^T,   z_j = U Q y_j,
    // the same spectral function as the reference's V f(Lambda) V^T rebuild
    // (project_to_PD.cpp:13-32). Elements whose block rows do not sum to zero, or with more
    // than three eigenvalues of M below eps, are handed to the workgroup-per-matrix kernel.

    // An element Hessian counts as translation invariant when each block-row sum is below
    // this fraction of max|H| (per block for the large kernel): round-off of FP64 sums, or the
    // rounding of FP32 storage (2^-24 per entry, a few entries per sum). The dropped coupling
    // with the translations is then at the level of the storage rounding.
    template <typename T>
    __host__ __device__ constexpr double translation_tol()
    {
        return sizeof(T) == 4 ? 1e-6 : 1e-10;
    }

    // Eigenvalues below this value are projected: eps (project_to_PD.cpp:13-32). With FP32
    // storage the stored N x N matrix carries rounding of up to 2^-24 max|H| per entry, which
    // moves its eigenvalues by up to N 2^-24 max|H| (Frobenius bound); eigenvalues in
    // [-N 2^-24 max|H|, eps) are then indistinguishable from that rounding, and so is their
    // projection (a change of H below the storage's own resolution), so they are left alone.
    template <typename T, int N>
    __device__ __forceinline__ double projection_cut(double eps, double scale)
    {
        if constexpr (sizeof(T) == 4) return fmin(eps, -(double)N * 0x1p-24 * scale);
        else return eps;
    }

    // Block (a, b), a <= b, of upper-block index u (inverse of ub4).
    __host__ __device__ constexpr int ub4_row(int u) { return u < 4 ? 0 : (u < 7 ? 1 : (u < 9 ? 2 : 3)); }
    __host__ __device__ constexpr int ub4_col(int u) { return u < 4 ? u : (u < 7 ? u - 3 : (u < 9 ? u - 5 : 3)); }

    // Stored value T (block u = T / 9, entry (i, j)) is H[(a,i),(b,j)] and, for a < b, also
    // H[(b,j),(a,i)]. It is added to the block-row sums rs[9 A + 3 I + J] = sum_B
    // H[(A,I),(B,J)], the scale max|H| and the upper triangle of M(al,I; be,J) =
    // sum_{A,B} w(A,al) w(B,be) H[(A,I),(B,J)]; the values may arrive in any order.
    template <int T>
    __device__ __forceinline__ void r4_take_value(double v, double (&m)[Sym<9>::size], double (&rs)[36], double& scale)
    {
        constexpr int u = T / 9, i = (T % 9) / 3, j = T % 3, a = ub4_row(u), b = ub4_col(u);
        scale = fmax(scale, fabs(v));
        rs[9 * a + 3 * i + j] += v;
        if constexpr (a != b) rs[9 * b + 3 * j + i] += v;
        static_for<0, 3>([&](auto Al) {
            static_for<decltype(Al)::value, 3>([&](auto Be) {
                constexpr int al = decltype(Al)::value, be = decltype(Be)::value;
                if constexpr (al != be || i <= j) {
                    constexpr double w = helmert(a, al) * helmert(b, be);
                    if constexpr (w != 0.0) m[Sym<9>::idx(3 * al + i, 3 * be + j)] += w * v;
                }
                if constexpr (a != b && (al != be || j <= i)) {
                    constexpr double w = helmert(b, al) * helmert(a, be);
                    if constexpr (w != 0.0) m[Sym<9>::idx(3 * al + j, 3 * be + i)] += w * v;
                }
            });
        });
    }

    // Eigenpairs of M below eps: the eigenvectors of M in y (Q y_S; the update vectors are
    // z_S = U y_S) with weights dl_S (zero for S >= cnt). m is overwritten. Returns false
    // when the element must take the workgroup path (more than three eigenvalues below eps,
    // or an eigenvalue that left its Sturm bracket). hv: 28 lane-interleaved LDS doubles
    // (stride 64) for the reflectors.
    __device__ __forceinline__ bool r4_eigen_update(double (&m)[Sym<9>::size], double eps, double cut, bool mirror, double* hv,
                                                    double (&y)[3][9], double (&dl)[3])
    {
        double d[9], e[9], tau[9];
        tridiag_reg<9>(m, d, e, tau);
        // the reflectors (28 entries) wait in LDS (hv: 28 x 64 doubles, lane-interleaved)
        // for the back-transform, which keeps the inverse iteration within 256 VGPRs
        static_for<0, 7>([&](auto Kc) {
            constexpr int K = decltype(Kc)::value, B = 7 * K - (K * (K - 1)) / 2;  // sum_{k<K} (7 - k)
            static_for<K + 2, 9>([&](auto Ic) { hv[(B + Ic() - K - 2) * 64] = m[Sym<9>::idx(K, Ic())]; });
        });
        asm volatile("" ::: "memory");  // reload them from LDS instead of keeping the registers
        double tn = 0.0, glo = 0.0, emax = 0.0;
        static_for<0, 9>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            const double l = I > 0 ? fabs(e[I > 0 ? I - 1 : 0]) : 0.0, r = fabs(e[I]);
            tn = fmax(tn, fabs(d[I]) + l + r);
            glo = I == 0 ? d[0] - l - r : fmin(glo, d[I] - l - r);
            emax = fmax(emax, r * r);
        });
        const double ulp = 2.220446049250313e-16;
        const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
        const double tol = 4.0 * ulp * tn;
        glo -= 2.0 * ulp * tn + pivmin;
        // T scaled by a power of two (exactly) to ||T|| < 1 for the minor recurrence of the
        // Sturm counts (sturm3_det), whose steps are one FMA deep
        double ds[9], e2s[9];
        int tex = 0;
        frexp(tn, &tex);
        const double sc = tn > 0.0 ? ldexp(1.0, -tex) : 1.0;
        static_for<0, 9>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            ds[I] = d[I] * sc;
            const double es = e[I] * sc;
            e2s[I] = es * es;
        });
        int cnt;
        {
            const double xe[3] = { cut * sc, cut * sc, cut * sc };
            int c3[3];
            sturm3_det<9>(ds, e2s, xe, c3);
            cnt = c3[0];
        }
        if (cnt > 3) return false;
        // brackets of the cnt lowest eigenvalues by bisection (three in step) down to about
        // 1e-7 ||T||; Rayleigh-quotient iteration below finishes them
        double lo[3] = { glo, glo, glo }, hi[3] = { cut, cut, cut };
        for (int it = 0; it < 24; it++) {
            bool busy = false;
            static_for<0, 3>([&](auto Jc) { busy |= Jc() < cnt && hi[Jc()] - lo[Jc()] > tol; });
            if (!busy) break;
            double mid[3], mids[3];
            int c3[3];
            static_for<0, 3>([&](auto Jc) {
                mid[Jc()] = 0.5 * (lo[Jc()] + hi[Jc()]);
                mids[Jc()] = mid[Jc()] * sc;
            });
            sturm3_det<9>(ds, e2s, mids, c3);
            static_for<0, 3>([&](auto Jc) {
                constexpr int J = decltype(Jc)::value;
                if (hi[J] - lo[J] > tol) {
                    if (c3[J] > J) hi[J] = mid[J];
                    else lo[J] = mid[J];
                }
            });
        }
        double lam[3];
        static_for<0, 3>([&](auto Jc) { lam[Jc()] = 0.5 * (lo[Jc()] + hi[Jc()]); });
        // eigenpairs of T by Rayleigh-quotient iteration from the brackets (dstein order):
        // slot by slot, three solves with the shift updated to y^T T y after each, every solve
        // followed by Gram-Schmidt against the finished lower slots, which separates
        // numerically degenerate eigenvalues (e.g. the three rotations of an unstressed
        // element). The converged quotient must stay in its Sturm bracket (widened by the
        // bracket width), else the element takes the Jacobi path.
        const double ptol = ulp * tn, bigr = 1.0 / ptol;
        bool ok = true;
        static_for<0, 3>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            static_for<0, 9>([&](auto Ic) { y[J][Ic()] = start_component(J, Ic()); });
            double nn = 1.0;
            if (J < cnt) {
                double sigma = lam[J];
                for (int it = 0; it < 3; it++) {
                    double ymax = 0.0;
                    static_for<0, 9>([&](auto Ic) { ymax = fmax(ymax, fabs(y[J][Ic()])); });
                    const double sc = 1.0 / ymax;
                    static_for<0, 9>([&](auto Ic) { y[J][Ic()] *= sc; });
                    tri_solve_reg<9>(d, e, sigma, ptol, bigr, y[J]);
                    static_for<0, J>([&](auto Pc) {
                        double dp = 0.0;
                        static_for<0, 9>([&](auto Ic) { dp += y[Pc()][Ic()] * y[J][Ic()]; });
                        static_for<0, 9>([&](auto Ic) { y[J][Ic()] -= dp * y[Pc()][Ic()]; });
                    });
                    double num = 0.0;
                    nn = 0.0;
                    static_for<0, 9>([&](auto Ic) {
                        constexpr int I = decltype(Ic)::value;
                        double ty = d[I] * y[J][I];
                        if constexpr (I > 0) ty += e[I - 1] * y[J][I - 1];
                        if constexpr (I < 8) ty += e[I] * y[J][I + 1];
                        num += y[J][I] * ty;
                        nn += y[J][I] * y[J][I];
                    });
                    sigma = num / nn;
                }
                const double w = hi[J] - lo[J] + tol;
                ok = ok && sigma >= lo[J] - w && sigma <= hi[J] + w;
                lam[J] = sigma;
            }
            const double inv = J < cnt ? 1.0 / sqrt(nn) : 0.0;
            static_for<0, 9>([&](auto Ic) { y[J][Ic()] *= inv; });
        });
        if (!ok) return false;
        // back-transform y <- H_0 ... H_6 y
        static_for<0, 3>([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            static_for<0, 7>([&](auto Kc) {
                constexpr int K = 6 - decltype(Kc)::value, B = 7 * K - (K * (K - 1)) / 2;
                double dp = y[J][K + 1];
                static_for<K + 2, 9>([&](auto Ic) { dp += hv[(B + Ic() - K - 2) * 64] * y[J][Ic()]; });
                dp *= tau[K];
                y[J][K + 1] -= dp;
                static_for<K + 2, 9>([&](auto Ic) { y[J][Ic()] -= dp * hv[(B + Ic() - K - 2) * 64]; });
            });
            dl[J] = J < cnt ? (mirror ? -2.0 * lam[J] : eps - lam[J]) : 0.0;
        });
        return true;
    }

    // One wave's 64 consecutive 90-value records, values [c0, c0 + W): global <-> LDS
    // stage[rec * SS + k] with consecutive lanes on consecutive addresses (whole runs of a
    // record per instruction, instead of 64 cache lines when each lane reads its own record).
    // Loads cover the records below nv (nv >= 1); stores only those with use[rec] != 0.
    template <int W, int SS, typename T>
    __device__ __forceinline__ void r4_stage_in(const T* __restrict__ base, int64_t rs, int nv, double* stage, int lane)
    {
        T v[W];
        if constexpr (sizeof(T) == 4) {
            // Every lane loads (records past nv read the last valid record and are zeroed
            // afterwards): a conditional load becomes a branch per load with the wait, and
            // the widening to FP64, inside it, which serialises the loads. (For FP64 storage
            // the conditional form schedules better within the 256-VGPR budget.)
#pragma unroll
            for (int s = 0; s < W; s++) {
                const int k = lane + 64 * s, rec = k / W, q = k - rec * W;
                v[s] = base[(int64_t)(rec < nv ? rec : nv - 1) * rs + q];
            }
        }
        else {
#pragma unroll
            for (int s = 0; s < W; s++) {
                const int k = lane + 64 * s, rec = k / W, q = k - rec * W;
                v[s] = rec < nv ? base[(int64_t)rec * rs + q] : T(0);
            }
        }
#pragma unroll
        for (int s = 0; s < W; s++) {
            const int k = lane + 64 * s, rec = k / W, q = k - rec * W;
            stage[rec * SS + q] = rec < nv ? (double)v[s] : 0.0;
        }
    }
    template <int W, int SS, typename T>
    __device__ __forceinline__ void r4_stage_out(T* __restrict__ base, int64_t rs, const int* use, const double* stage, int lane)
    {
#pragma unroll 4
        for (int k = lane; k < 64 * W; k += 64) {
            const int rec = k / W, q = k - rec * W;
            if (use[rec]) base[(int64_t)rec * rs + q] = (T)stage[rec * SS + q];
        }
    }

    __device__ inline void count_changed(bool changed, unsigned long long* count)
    {
        if (count != nullptr) {
            const unsigned long long m = __ballot(changed);
            if ((threadIdx.x & 63) == 0 && m != 0ull) atomicAdd(count, (unsigned long long)__popcll(m));
        }
    }

    // 4-block elements, one element per lane, 64 per workgroup (one wave): reduced 9x9 path;
    // the elements it hands back are flagged in needs_full. The element records are read and
    // written through LDS 16 values at a time with coalesced transfers: M, the block-row
    // sums and the scale accumulate chunk by chunk, and the update is added chunk by chunk.
    template <typename TH>
    __global__ __launch_bounds__(64, 2) void k_project_ub4_reduced(TH* __restrict__ H, int64_t n_mats, int64_t rec_stride, double eps, int mirror,
                                                               unsigned long long* count, uint8_t* __restrict__ needs_full,
                                                               const uint8_t* __restrict__ select)
    {
        // reflectors during the eigen-solve (28 x 64, lane-interleaved); staging of the
        // element records (64 x 17) before and after it
        // Chunks of CW values per record: 16 doubles or 32 floats (128-byte runs either way)
        constexpr int CW = sizeof(TH) == 4 ? 32 : 16, SS = CW + 1, NCH = (90 + CW - 1) / CW;
        __shared__ double s_lds[(64 * SS > 28 * 64) ? 64 * SS : 28 * 64];
        __shared__ int s_use[64];
        const int lane = threadIdx.x;
        const int64_t e0 = (int64_t)blockIdx.x * 64, e = e0 + lane;
        const bool active = e < n_mats && (select == nullptr || select[e]);
        const int nv = n_mats - e0 < 64 ? (int)(n_mats - e0) : 64;  // records in this wave
        TH* const base = H + e0 * rec_stride;
        s_use[lane] = active ? 1 : 0;
        __syncthreads();
        bool ok = false;
        double y[3][9], dl[3];
        if (__ballot(active) != 0ull) {
            double m[Sym<9>::size], rs[36], scale = 0.0;
            static_for<0, Sym<9>::size>([&](auto Kc) { m[Kc()] = 0.0; });
            static_for<0, 36>([&](auto Kc) { rs[Kc()] = 0.0; });
            static_for<0, NCH>([&](auto Cc) {
                constexpr int c0 = CW * decltype(Cc)::value, W = 90 - c0 < CW ? 90 - c0 : CW;
                r4_stage_in<W, SS>(base + c0, rec_stride, nv, s_lds, lane);
                __syncthreads();
                static_for<0, W>([&](auto Kc) { r4_take_value<c0 + decltype(Kc)::value>(s_lds[lane * SS + Kc()], m, rs, scale); });
                __syncthreads();
            });
            if (active) {
                double rowsum = 0.0;
                static_for<0, 36>([&](auto Kc) { rowsum = fmax(rowsum, fabs(rs[Kc()])); });
                ok = rowsum <= translation_tol<TH>() * scale && r4_eigen_update(m, eps, projection_cut<TH, 12>(eps, scale), mirror != 0, s_lds + lane, y, dl);
            }
        }
        if (e < n_mats) needs_full[e] = (active && !ok) ? 1 : 0;
        __syncthreads();  // every lane is done with the reflectors
        s_use[lane] = ok ? 1 : 0;
        __syncthreads();
        if (__ballot(ok) != 0ull) {
            // f(0) on span{translations}: eps T T^T, T T^T = (1/4) 1 1^T (x) I3
            const double null_term = mirror ? 0.0 : 0.25 * eps;
            // z_S = U y_S, U = W (x) I3
            double z[3][12];
            static_for<0, 3>([&](auto Sc) {
                static_for<0, 4>([&](auto Ac) {
                    static_for<0, 3>([&](auto Ic) {
                        constexpr int S = decltype(Sc)::value, A = decltype(Ac)::value, I = decltype(Ic)::value;
                        double acc = 0.0;
                        static_for<0, 3>([&](auto Al) {
                            constexpr double w = helmert(A, decltype(Al)::value);
                            if constexpr (w != 0.0) acc += w * y[S][3 * decltype(Al)::value + I];
                        });
                        z[S][3 * A + I] = acc;
                    });
                });
            });
            static_for<0, NCH>([&](auto Cc) {
                constexpr int c0 = CW * decltype(Cc)::value, W = 90 - c0 < CW ? 90 - c0 : CW;
                r4_stage_in<W, SS>(base + c0, rec_stride, nv, s_lds, lane);
                __syncthreads();
                if (ok) {
                    static_for<0, W>([&](auto Kc) {
                        constexpr int T = c0 + decltype(Kc)::value, u = T / 9, i = (T % 9) / 3, j = T % 3;
                        constexpr int a = ub4_row(u), b = ub4_col(u);
                        double acc = (i == j) ? null_term : 0.0;
                        static_for<0, 3>([&](auto Sc) {
                            constexpr int S = decltype(Sc)::value;
                            acc += dl[S] * z[S][3 * a + i] * z[S][3 * b + j];
                        });
                        s_lds[lane * SS + Kc()] += acc;
                    });
                }
                __syncthreads();
                r4_stage_out<W, SS>(base + c0, rec_stride, s_use, s_lds, lane);
                __syncthreads();
            });
        }
        count_changed(ok, count);
    }

    template <int NB, typename T>
    __global__ __launch_bounds__(64) void k_project_ub_full(T* __restrict__ H, int64_t n_mats, int64_t rs, double eps, int mirror,
                                                           unsigned long long* count, const uint8_t* __restrict__ only)
    {
        constexpr int NUB = NB * (NB + 1) / 2;
        const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
        bool changed = false;
        if (e < n_mats && (only == nullptr || only[e])) changed = project_element_full<NB>(H + e * rs, eps, mirror != 0);
        count_changed(changed, count);
    }

    // Large element Hessians (8 and more DoF blocks: Hex8, Tet10, Hex27) and the dense
    // standalone entry for n > 12: one workgroup per matrix, the matrix in LDS. Instead of
    // a full eigen-decomposition (project_to_PD.cpp:13-32 rebuilds V f(Lambda) V^T) only the
    // eigenpairs below eps are computed, and the projection is written as the rank-k update
    //     H + sum_j (f(lambda_j) - lambda_j) u_j u_j^T,   f = eps (clamp) or -lambda (mirror),
    // which is the same spectral function. Steps:
    //  1. Translation-invariant element matrices (every block row sums to zero, as for every
    //     elastic energy) get the translations shifted up by max|H| (adding max|H|/NB to the
    //     same-component entries of every 3x3 block), which leaves every other eigenpair
    //     alone; their exact zero eigenvalues then map to f(0) directly: + eps/NB (1 1^T (x) I3)
    //     for clamp, nothing for mirror, as the register kernels of 4-node elements do.
    //  2. Householder tridiagonalisation T = Q^T A Q (LAPACK dsytd2 arithmetic; v_0 = 1 and
    //     the rest of v stored below the subdiagonal). Two workgroup barriers per column.
    //  3. Sturm count of T - eps I: the number k of eigenvalues below eps. k = 0 ends here.
    //  4. The k eigenvalues by multisection: every thread evaluates Sturm counts at two
    //     points per round; the first round spreads all points over [Gershgorin low, eps],
    //     later rounds split each eigenvalue's bracket, until every bracket is below
    //     4 ulp(||T||).
    //  5. Eigenvectors of T by inverse iteration (dstein): LU of T - lambda I with partial
    //     pivoting, three solves from a fixed pseudo-random start, tiny pivots replaced by
    //     ulp(||T||); one lane per eigenvalue, Gram-Schmidt within clusters of close
    //     eigenvalues between the solves.
    //  6. Back-transform u = Q y (one wave per vector, no barriers), then the rank-k update.
    constexpr int PNT = 256;   // threads per matrix
    constexpr int PPTS = 2;    // Sturm points per thread per round
    constexpr int PLU = 16;    // lanes running inverse iteration at once (full-size variant)

    // Full-wave reductions through DPP (quad permutes and row rotations stay in the VALU;
    // __shfl_xor goes through ds_bpermute at LDS latency), then the four row results read
    // as scalars. Every lane must be active; the result is the same in every lane and the
    // summation order is fixed.
    template <int CTRL>
    __device__ __forceinline__ double dpp_mov(double v)
    {
        // quad permutes and row rotations read a valid lane everywhere, so no "old" value is
        // needed (update_dpp with old = 0 would cost two extra moves per call)
        const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }
    __device__ __forceinline__ double readlane_d(double v, int l)
    {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
    }
    __device__ inline double wave_sum_all(double v)
    {
        v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
        v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
        v += dpp_mov<0x124>(v);  // row_ror:4
        v += dpp_mov<0x128>(v);  // row_ror:8
        return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
    }
    __device__ inline double wave_max_all(double v)
    {
        v = fmax(v, dpp_mov<0xB1>(v));
        v = fmax(v, dpp_mov<0x4E>(v));
        v = fmax(v, dpp_mov<0x124>(v));
        v = fmax(v, dpp_mov<0x128>(v));
        return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
    }
    __device__ inline double wave_min_all(double v)
    {
        v = fmin(v, dpp_mov<0xB1>(v));
        v = fmin(v, dpp_mov<0x4E>(v));
        v = fmin(v, dpp_mov<0x124>(v));
        v = fmin(v, dpp_mov<0x128>(v));
        return fmin(fmin(readlane_d(v, 0), readlane_d(v, 16)), fmin(readlane_d(v, 32), readlane_d(v, 48)));
    }
    // Number of eigenvalues of the tridiagonal (dg, od) below x0 and below x1 (two
    // interleaved recurrences; a pivot smaller than pivmin counts as -pivmin, dlaneg).
    template <int N>
    __device__ inline void sturm2(const double* dg, const double* od, double x0, double x1, double pivmin, int& c0,
                                  int& c1)
    {
        double q0 = dg[0] - x0, q1 = dg[0] - x1;
        if (fabs(q0) < pivmin) q0 = -pivmin;
        if (fabs(q1) < pivmin) q1 = -pivmin;
        int n0 = q0 < 0.0, n1 = q1 < 0.0;
        for (int i = 1; i < N; i++) {
            const double d = dg[i], e = od[i - 1], e2 = e * e;
            q0 = (d - x0) - e2 * rcp_nr1(q0);  // |q| >= pivmin keeps e2 / q finite
            q1 = (d - x1) - e2 * rcp_nr1(q1);
            if (fabs(q0) < pivmin) q0 = -pivmin;
            if (fabs(q1) < pivmin) q1 = -pivmin;
            n0 += q0 < 0.0;
            n1 += q1 < 0.0;
        }
        c0 = n0;
        c1 = n1;
    }

    // Number of eigenvalues of the tridiagonal (dg, od) below x (sturm2 with one point).
    template <int N>
    __device__ inline int sturm1(const double* dg, const double* od, double x, double pivmin)
    {
        double q = dg[0] - x;
        if (fabs(q) < pivmin) q = -pivmin;
        int n = q < 0.0;
        for (int i = 1; i < N; i++) {
            const double e = od[i - 1];
            q = (dg[i] - x) - e * e * rcp_nr1(q);  // |q| >= pivmin keeps e2 / q finite
            if (fabs(q) < pivmin) q = -pivmin;
            n += q < 0.0;
        }
        return n;
    }

    // LDS of the phases after the tridiagonalisation (count, multisection, inverse iteration,
    // back-transform, update)
    template <int N, int PT, int KMAX, int PL>
    struct ProjTail
    {
        double Y[KMAX * N];
        double dg[N], od[N], tau[N], lamv[N], bra[N], brb[N];
        double lu[3 * N * PL];
        double lm[N * PL];
        uint8_t lsw[N * PL];
        int cbuf[PT * PPTS];
        int busy[3];
    };

    // Steps 3-7 of the large projection on the tridiagonal T = Q^T A Q in L.dg / L.od (and
    // L.tau, written and synchronised by the caller). vref(i, k) is entry i >= k + 2 of the
    // Householder vector of column k (v_{k+1} = 1, zero above). Every early return is uniform;
    // true: more than KMAX eigenvalues below eps, the matrix is left untouched.
    template <int N, bool DENSE, int PT, typename T, int KMAX, int PL, typename VREF>
    __device__ bool project_tail(T* __restrict__ h, int64_t e, double eps, int mirror, unsigned long long* count,
                                 uint8_t* __restrict__ changed_out, bool ti, double scale, ProjTail<N, PT, KMAX, PL>& L, VREF vref)
    {
        constexpr int NB = N / 3, NUB = NB * (NB + 1) / 2, NW = PT / 64, NPT = PT * PPTS;
        const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
        double* Y = L.Y;
        double* dg = L.dg;
        double* od = L.od;
        double* tau = L.tau;
        double* lamv = L.lamv;
        double* bra = L.bra;
        double* brb = L.brb;
        double* lu = L.lu;
        int* cbuf = L.cbuf;
        int* busy = L.busy;
        auto ubi = [](int a, int b) { return a * NB - (a * (a - 1)) / 2 + (b - a); };
        // element (r, c) of the upper-block storage, r <= c blockwise
        auto ub_index = [&](int t, int& r, int& c) {
            const int u = t / 9, ij = t - 9 * u;
            int a = 0, rem = u;
            while (rem >= NB - a) {
                rem -= NB - a;
                a++;
            }
            r = 3 * a + ij / 3;
            c = 3 * (a + rem) + ij % 3;
        };
        // ---- 3. count below eps -----------------------------------------------------------
        double tn = 0.0, glo = 0.0, emax = 0.0;
        {
            double tnl = 0.0, gl = INFINITY, el = 0.0;
            for (int i = lane; i < N; i += 64) {
                const double l = i > 0 ? fabs(od[i - 1]) : 0.0, r = i < N - 1 ? fabs(od[i]) : 0.0;
                tnl = fmax(tnl, fabs(dg[i]) + l + r);
                gl = fmin(gl, dg[i] - l - r);
                el = fmax(el, r * r);
            }
            tn = wave_max_all(tnl);
            glo = wave_min_all(gl);
            emax = wave_max_all(el);
        }
        const double ulp = 2.220446049250313e-16;
        const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
        const double tol = 4.0 * ulp * tn;
        glo -= 2.0 * ulp * tn + pivmin;
        const double cut = DENSE ? eps : projection_cut<T, N>(eps, scale);
        int cnt, dummy;
        sturm2<N>(dg, od, cut, cut, pivmin, cnt, dummy);
        if (cnt == 0) {
            if (ti && !mirror) {
                for (int t = tid; t < NUB * 9; t += PT) {
                    const int ij = t % 9;
                    if (ij / 3 == ij % 3) h[t] += eps / NB;
                }
                if (tid == 0 && count != nullptr) atomicAdd(count, 1ull);
            }
            if (tid == 0 && changed_out != nullptr) changed_out[e] = (ti && !mirror) ? 1 : 0;
            return false;
        }
        if (cnt > KMAX) return true;
        // ---- 4. multisection for the cnt lowest eigenvalues ------------------------------
        // A round costs the instructions of every wave that holds a point. Up to 8 eigenvalues
        // (the rule for element Hessians) take all points of a round in wave 0, one per lane
        // (64 / cnt per eigenvalue, at least 3.2 bits a round): fewer wave-rounds than more
        // points over more waves. Larger counts take two points per thread over the workgroup.
        {
            const bool one_wave = cnt <= 8;
            const int NP0 = one_wave ? 64 : NPT;
            const double w0 = (cut - glo) / (NP0 + 1);
            if (one_wave) {
                if (tid < 64) cbuf[tid] = sturm1<N>(dg, od, glo + w0 * (tid + 1), pivmin);
            }
            else {
                const int p0 = tid * PPTS;
                int c0, c1;
                sturm2<N>(dg, od, glo + w0 * (p0 + 1), glo + w0 * (p0 + 2), pivmin, c0, c1);
                cbuf[p0] = c0;
                cbuf[p0 + 1] = c1;
            }
            if (tid == 0) busy[0] = 0;
            __syncthreads();
            for (int j = tid; j < cnt; j += PT) {  // first point with more than j eigenvalues below it
                int lo = 0, hi = NP0;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cbuf[mid] > j) hi = mid;
                    else lo = mid + 1;
                }
                bra[j] = lo > 0 ? glo + w0 * lo : glo;
                brb[j] = lo < NP0 ? glo + w0 * (lo + 1) : cut;
            }
            __syncthreads();
            // later rounds: P points per eigenvalue still above tol
            const int P = min(64, (one_wave ? 64 : NPT) / cnt);
            for (int round = 0; round < 20; round++) {
                if (one_wave) {
                    const int j = tid / P, l = tid - j * P;
                    if (tid < 64 && j < cnt && brb[j] - bra[j] > tol)
                        cbuf[tid] = sturm1<N>(dg, od, bra[j] + (brb[j] - bra[j]) * ((double)(l + 1) / (P + 1)), pivmin);
                }
                else {
                    int pts[PPTS];
                    double xs[PPTS];
#pragma unroll
                    for (int q = 0; q < PPTS; q++) {
                        const int p = tid * PPTS + q, j = p / P, l = p - j * P;
                        pts[q] = j < cnt ? p : -1;
                        xs[q] = j < cnt ? bra[j] + (brb[j] - bra[j]) * ((double)(l + 1) / (P + 1)) : 0.0;
                    }
                    if (pts[0] >= 0) {
                        int c0, c1;
                        sturm2<N>(dg, od, xs[0], xs[1], pivmin, c0, c1);
                        cbuf[pts[0]] = c0;
                        if (pts[1] >= 0) cbuf[pts[1]] = c1;
                    }
                }
                if (tid == 0) busy[(round + 1) % 3] = 0;
                __syncthreads();
                for (int j = tid; j < cnt; j += PT) {
                    const double a = bra[j], b = brb[j];
                    if (b - a > tol) {
                        int lo = 0, hi = P;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (cbuf[j * P + mid] > j) hi = mid;
                            else lo = mid + 1;
                        }
                        const double na = lo > 0 ? a + (b - a) * ((double)lo / (P + 1)) : a;
                        const double nb = lo < P ? a + (b - a) * ((double)(lo + 1) / (P + 1)) : b;
                        bra[j] = na;
                        brb[j] = nb;
                        if (nb - na > tol) busy[round % 3] = 1;
                    }
                }
                __syncthreads();
                if (!busy[round % 3]) break;
            }
            for (int j = tid; j < cnt; j += PT) lamv[j] = 0.5 * (bra[j] + brb[j]);
            __syncthreads();
        }
        // ---- 5. block inverse iteration (dstein), one lane per eigenvalue --------------------
        // Three rounds: every vector is solved with its own shift (LU with partial pivoting
        // fused with the forward substitution; U's reciprocal pivots and two superdiagonals in
        // LDS), then wave 0 normalises the vectors and orthonormalises those whose eigenvalues
        // are within 1e-9 ||T|| of each other (numerically degenerate: inverse iteration alone
        // cannot separate them). A last pass orthonormalises chains of eigenvalues within
        // 1e-3 ||T|| (dstein's reorthogonalisation threshold).
        {
            double* ur = lu;
            double* u1 = lu + N * PL;
            double* u2 = lu + 2 * N * PL;
            double* lm = L.lm;
            uint8_t* lsw = L.lsw;
            const double ptol = ulp * tn, bigr = 1.0 / ptol;
            for (int t = tid; t < cnt * N; t += PT) {
                const int j = t / N;
                Y[t] = start_component(j, t - j * N);
            }
            __syncthreads();
            auto orthonormalise = [&](double gtol, int passes) {  // wave 0
                int j0 = 0;
                for (int j = 0; j < cnt; j++) {
                    if (j == 0 || lamv[j] - lamv[j - 1] > gtol) j0 = j;
                    double* y = Y + j * N;
                    double y0 = lane < N ? y[lane] : 0.0, y1 = lane + 64 < N ? y[lane + 64] : 0.0;
                    for (int pass = 0; pass < (j > j0 ? passes : 0); pass++) {
                        const double inv = 1.0 / sqrt(wave_sum_all(y0 * y0 + y1 * y1));
                        y0 *= inv;
                        y1 *= inv;
                        for (int i = j0; i < j; i++) {
                            const double* u = Y + i * N;
                            const double v0 = lane < N ? u[lane] : 0.0, v1 = lane + 64 < N ? u[lane + 64] : 0.0;
                            const double d = wave_sum_all(v0 * y0 + v1 * y1);
                            y0 -= d * v0;
                            y1 -= d * v1;
                        }
                    }
                    const double inv = 1.0 / sqrt(wave_sum_all(y0 * y0 + y1 * y1));
                    if (lane < N) y[lane] = y0 * inv;
                    if (lane + 64 < N) y[lane + 64] = y1 * inv;
                }
            };
            for (int it = 0; it < 3; it++) {
                if (tid < PL) {
                    // with at most PL vectors each lane keeps its factors of T - lambda I (the
                    // shift does not change) and later rounds replay the row interchanges and
                    // multipliers of the forward pass instead of factoring again
                    const bool replay = it > 0 && cnt <= PL;
                    for (int j = tid; j < cnt; j += PL) {
                        double* y = Y + j * N;
                        double r0 = y[0];
                        if (replay) {
                            for (int k = 0; k < N - 1; k++) {
                                const double r1 = y[k + 1], mult = lm[k * PL + tid];
                                const bool sw = lsw[k * PL + tid] != 0;
                                y[k] = sw ? r1 : r0;
                                r0 = sw ? r0 - mult * r1 : r1 - mult * r0;
                            }
                        }
                        else {
                            const double lam = lamv[j];
                            double a = dg[0] - lam, c = od[0];
                            for (int k = 0; k < N - 1; k++) {
                                const double b = od[k], an = dg[k + 1] - lam, cn = k + 1 < N - 1 ? od[k + 1] : 0.0;
                                const double r1 = y[k + 1];
                                // partial pivoting: rows k and k + 1 interchanged when |b| > |a|
                                const bool sw = fabs(a) < fabs(b);
                                const double piv = sw ? b : a, oth = sw ? a : b;
                                const double inv = fabs(piv) >= 2.2250738585072014e-308 ? rcp_nr(piv) : 0.0;
                                const double mult = oth * inv;  // |mult| <= 1; a = b = 0 gives 0
                                lm[k * PL + tid] = mult;
                                lsw[k * PL + tid] = sw ? 1 : 0;
                                u1[k * PL + tid] = sw ? an : c;
                                u2[k * PL + tid] = sw ? cn : 0.0;
                                y[k] = sw ? r1 : r0;
                                r0 = sw ? r0 - mult * r1 : r1 - mult * r0;
                                const double an2 = sw ? c - mult * an : an - mult * c;
                                c = sw ? -mult * cn : cn;
                                a = an2;
                                ur[k * PL + tid] = fabs(piv) < ptol ? (piv < 0.0 ? -bigr : bigr) : inv;
                            }
                            ur[(N - 1) * PL + tid] = fabs(a) < ptol ? (a < 0.0 ? -bigr : bigr) : rcp_nr(a);
                        }
                        double y1 = r0 * ur[(N - 1) * PL + tid], y2 = 0.0;
                        y[N - 1] = y1;
                        for (int k = N - 2; k >= 0; k--) {  // back substitution, U has two superdiagonals
                            const double v = (y[k] - u1[k * PL + tid] * y1 - u2[k * PL + tid] * y2) * ur[k * PL + tid];
                            y[k] = v;
                            y2 = y1;
                            y1 = v;
                        }
                    }
                }
                __syncthreads();
                if (wid == 0) orthonormalise(1e-9 * tn, 2);
                __syncthreads();
            }
            if (wid == 0) orthonormalise(1e-3 * tn, 1);
            __syncthreads();
        }
        // ---- 6. back-transform u = H_0 H_1 ... H_{N-3} y, one wave per vector -----------------
        for (int j = wid; j < cnt; j += NW) {
            double* y = Y + j * N;
            double y0 = lane < N ? y[lane] : 0.0, y1 = lane + 64 < N ? y[lane + 64] : 0.0;
            for (int k = N - 3; k >= 0; k--) {
                const double t = tau[k];
                if (t == 0.0) continue;
                // v: 0 above k + 1, 1 at k + 1, vref below
                const double v0 = lane < k + 1 || lane >= N ? 0.0 : (lane == k + 1 ? 1.0 : vref(lane, k));
                const int i1 = lane + 64;
                const double v1 = i1 < k + 1 || i1 >= N ? 0.0 : (i1 == k + 1 ? 1.0 : vref(i1, k));
                const double d = t * wave_sum_all(v0 * y0 + v1 * y1);
                y0 -= d * v0;
                y1 -= d * v1;
            }
            if (lane < N) y[lane] = y0;
            if (lane + 64 < N) y[lane + 64] = y1;
        }
        __syncthreads();
        // ---- 7. rank-cnt update -------------------------------------------------------------
        for (int j = tid; j < cnt; j += PT) lamv[j] = mirror ? -2.0 * lamv[j] : eps - lamv[j];
        __syncthreads();
        if (DENSE) {
            for (int t = tid; t < N * N; t += PT) {
                const int r = t / N, c = t - r * N;
                double acc = 0.0;
                for (int j = 0; j < cnt; j++) acc += lamv[j] * Y[j * N + r] * Y[j * N + c];
                h[t] = (T)((double)h[t] + acc);
            }
        }
        else {
            const double tshift = (ti && !mirror) ? eps / NB : 0.0;
            for (int t = tid; t < NUB * 9; t += PT) {
                int r, c;
                ub_index(t, r, c);
                double acc = (r % 3) == (c % 3) ? tshift : 0.0;
                for (int j = 0; j < cnt; j++) acc += lamv[j] * Y[j * N + r] * Y[j * N + c];
                h[t] = (T)((double)h[t] + acc);
            }
        }
        if (tid == 0) {
            if (count != nullptr) atomicAdd(count, 1ull);
            if (changed_out != nullptr) changed_out[e] = 1;
        }
        return false;
    }


    // One matrix h (element e) with the whole workgroup of PT threads (a multiple of 64);
    // every early return is uniform. KMAX bounds the eigenvectors held in LDS and PL the
    // lanes running inverse iteration at once: with more than KMAX eigenvalues below eps the
    // matrix is left untouched and true is returned (the caller hands it to a variant with
    // KMAX = N). A smaller LDS footprint lets a CU hold several matrices at once.
    template <int N, bool DENSE, int PT, typename T, int KMAX = N, int PL = PLU>
    __device__ bool project_large_one(T* __restrict__ h, int64_t e, double eps, int mirror, unsigned long long* count,
                                      uint8_t* __restrict__ changed_out)
    {
        constexpr int NB = N / 3, LD = N | 1, NW = PT / 64;
        static_assert(N <= 128, "one wave holds a vector in two registers per lane");
        static_assert(KMAX >= 1 && KMAX <= N && PL >= 1 && PL <= 64, "eigenvector and solve lane bounds");
        __shared__ double A[N * LD];
        __shared__ double pv[N], wv[N];
        __shared__ double red[2 * NW];
        __shared__ ProjTail<N, PT, KMAX, PL> L;
        double* dg = L.dg;
        double* od = L.od;
        double* tau = L.tau;
        const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
        auto ubi = [](int a, int b) { return a * NB - (a * (a - 1)) / 2 + (b - a); };
        for (int t = tid; t < N * N; t += PT) {
            const int r = t / N, c = t - r * N;
            double v;
            if (DENSE) v = 0.5 * ((double)h[r * N + c] + (double)h[c * N + r]);
            else {
                const int a = r / 3, i = r - 3 * a, b = c / 3, j = c - 3 * b;
                v = a <= b ? (double)h[ubi(a, b) * 9 + 3 * i + j] : (double)h[ubi(b, a) * 9 + 3 * j + i];
            }
            A[r * LD + c] = v;
        }
        __syncthreads();
        auto block_max = [&](double v) {
            v = wave_max_all(v);
            if (lane == 0) red[wid] = v;
            __syncthreads();
            double r = red[0];
            for (int w = 1; w < NW; w++) r = fmax(r, red[w]);
            __syncthreads();
            return r;
        };
        // ---- 1. scale, translation invariance ---------------------------------------------
        double amax = 0.0, rsum = 0.0;
        for (int t = tid; t < N * N; t += PT) amax = fmax(amax, fabs(A[(t / N) * LD + t % N]));
        if (!DENSE) {
            for (int t = tid; t < N * 3; t += PT) {  // row r, component j: sum over blocks b of A[r][3b + j]
                const int r = t / 3, j = t - 3 * r;
                double acc = 0.0;
                for (int b = 0; b < NB; b++) acc += A[r * LD + 3 * b + j];
                rsum = fmax(rsum, fabs(acc));
            }
        }
        const double scale = block_max(amax);
        const bool ti = !DENSE && block_max(rsum) <= translation_tol<T>() * scale * NB && scale > 0.0;
        if (!(scale > 0.0)) {
            // zero matrix: every eigenvalue (0) is below eps -> eps I (clamp), 0 (mirror)
            if (!mirror) {
                for (int r = tid; r < N; r += PT) {
                    if (DENSE) h[r * N + r] = eps;
                    else h[ubi(r / 3, r / 3) * 9 + 4 * (r % 3)] = eps;
                }
            }
            if (tid == 0) {
                if (count != nullptr) atomicAdd(count, 1ull);
                if (changed_out != nullptr) changed_out[e] = 1;
            }
            return false;
        }
        if (ti) {
            const double sh = scale / NB;
            for (int t = tid; t < N * N; t += PT) {
                const int r = t / N, c = t - r * N;
                if ((r % 3) == (c % 3)) A[r * LD + c] += sh;
            }
            __syncthreads();
        }
        // ---- 2. Householder tridiagonalisation --------------------------------------------
        for (int k = 0; k < N - 2; k++) {
            const int m = N - k - 1;
            double s = 0.0;
            for (int i = k + 2 + lane; i < N; i += 64) {
                const double a = A[i * LD + k];
                s += a * a;
            }
            s = wave_sum_all(s);
            const double x0 = A[(k + 1) * LD + k];
            double beta = x0, t = 0.0, invd = 0.0;
            if (s != 0.0) {
                beta = -copysign(sqrt(x0 * x0 + s), x0);
                t = (beta - x0) / beta;
                invd = 1.0 / (x0 - beta);
            }
            if (tid < m) {  // p = t A22 v, column access (A22 symmetric)
                const int c = k + 1 + tid;
                double a0 = 0.0, a1 = 0.0;
                int j = 1;
                for (; j + 1 < m; j += 2) {
                    a0 += A[(k + 1 + j) * LD + c] * A[(k + 1 + j) * LD + k];
                    a1 += A[(k + 2 + j) * LD + c] * A[(k + 2 + j) * LD + k];
                }
                if (j < m) a0 += A[(k + 1 + j) * LD + c] * A[(k + 1 + j) * LD + k];
                pv[tid] = t * (A[(k + 1) * LD + c] + invd * (a0 + a1));
                wv[tid] = tid == 0 ? 1.0 : A[c * LD + k] * invd;
            }
            __syncthreads();
            double kk = 0.0;
            for (int i = lane; i < m; i += 64) kk += pv[i] * wv[i];
            kk = 0.5 * t * wave_sum_all(kk);
            if (t != 0.0) {
                const int step = PT / m;
                if (tid < step * m) {
                    const int j = tid % m;
                    const double vj = wv[j], wj = pv[j] - kk * vj;
                    for (int i = tid / m; i < m; i += step) {
                        const double vi = wv[i], wi = pv[i] - kk * vi;
                        A[(k + 1 + i) * LD + k + 1 + j] -= vi * wj + wi * vj;
                    }
                }
            }
            if (tid >= 1 && tid < m) A[(k + 1 + tid) * LD + k] = wv[tid];
            if (tid == 0) {
                dg[k] = A[k * LD + k];
                od[k] = beta;
                tau[k] = t;
            }
            __syncthreads();
        }
        if (tid == 0) {
            dg[N - 2] = A[(N - 2) * LD + N - 2];
            od[N - 2] = A[(N - 1) * LD + N - 2];
            dg[N - 1] = A[(N - 1) * LD + N - 1];
            od[N - 1] = 0.0;
            tau[N - 2] = 0.0;
            tau[N - 1] = 0.0;
        }
        __syncthreads();
        return project_tail<N, DENSE, PT, T, KMAX, PL>(h, e, eps, mirror, count, changed_out, ti, scale, L,
                                                        [&](int i, int k) { return A[i * LD + k]; });
    }

    // Sum over the 8 lanes of a half DPP row (quad swaps, then the half-row mirror): every
    // lane of the group gets the same, bitwise equal sum (each step adds two equal operands in
    // swapped order). All lanes must be active.
    __device__ inline double grp8_sum(double v)
    {
        v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
        v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
        v += dpp_mov<0x141>(v);  // row_half_mirror
        return v;
    }

    // project_large_one for the upper-block storage of large elements (Hex27: 81 x 81) with
    // the matrix in registers during the tridiagonalisation instead of in LDS. Thread
    // (ti, tj) of a 32 x 8 grid (tj = lane % 8, ti = 8 wave + lane / 8, so the 8 threads of a
    // half DPP row share ti) holds rows ti + 32 q (q < TR) and columns tj + 8 s (s < TC) in
    // a[q][s]:
    //  - A22 v is a partial product per thread summed over its 8 lanes (no barrier);
    //  - the step is specialised at compile time on the first column group s0 = (k + 1) / 8
    //    that still reaches the active block (and the row group q0 = s0 / 4), so the products
    //    and the rank-2 update shrink with the block;
    //  - two barriers per column (p, then the next column), against three LDS round trips per
    //    column entry in the LDS form; the Householder vectors go to a packed LDS array for
    //    the back-transform (25 KB instead of the 52 KB matrix), three workgroups per CU.
    // The kernel is bound by VALU issue, hence the narrow row groups: the per-column fixed
    // cost (two wave sums, the Householder scalars, selects) is paid by every wave, the
    // group sums cost 3 DPP steps per own row.
    // The arithmetic is that of project_large_one (dsytd2), summed in another order.
    template <int N, typename T, int KMAX, int PL>
    __device__ bool project_tiled_one(T* __restrict__ h, int64_t e, double eps, int mirror, unsigned long long* count,
                                      uint8_t* __restrict__ changed_out)
    {
        constexpr int PT = 256, NW = PT / 64, TI = 32, TJ = 8, TR = (N + TI - 1) / TI, TC = (N + TJ - 1) / TJ, NB = N / 3;
        constexpr int NV = (N - 1) * (N - 2) / 2;  // column k holds rows k + 2 .. N - 1
        static_assert(N % 3 == 0 && N <= 128 && TR * TI >= N && TC * TJ >= N && TI == 4 * TJ, "tile bounds");
        __shared__ double V[NV];
        __shared__ double xb[2][128];
        __shared__ double pb[128];
        __shared__ double red[2 * NW];
        __shared__ ProjTail<N, PT, KMAX, PL> L;
        const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
        const int ti = wid * 8 + (lane >> 3), tj = lane & 7;
        auto ubi = [](int a, int b) { return a * NB - (a * (a - 1)) / 2 + (b - a); };
        auto voff = [](int k) { return k * (N - 2) - (k * (k - 1)) / 2; };
        double a[TR][TC];
#pragma unroll
        for (int q = 0; q < TR; q++) {
#pragma unroll
            for (int s = 0; s < TC; s++) {
                const int r = ti + TI * q, c = tj + TJ * s;
                double v = 0.0;
                if (r < N && c < N) {
                    const int A_ = r / 3, i = r - 3 * A_, B_ = c / 3, j = c - 3 * B_;
                    v = A_ <= B_ ? (double)h[ubi(A_, B_) * 9 + 3 * i + j] : (double)h[ubi(B_, A_) * 9 + 3 * j + i];
                }
                a[q][s] = v;
            }
        }
        auto block_max = [&](double v) {
            v = wave_max_all(v);
            if (lane == 0) red[wid] = v;
            __syncthreads();
            double r = red[0];
            for (int w = 1; w < NW; w++) r = fmax(r, red[w]);
            __syncthreads();
            return r;
        };
        // ---- 1. scale, translation invariance (block-row sums per component) ---------------
        double amax = 0.0, rsum = 0.0;
#pragma unroll
        for (int q = 0; q < TR; q++) {
            double ps0 = 0.0, ps1 = 0.0, ps2 = 0.0;
#pragma unroll
            for (int s = 0; s < TC; s++) {
                const int cm = (tj + TJ * s) % 3;
                amax = fmax(amax, fabs(a[q][s]));
                ps0 += cm == 0 ? a[q][s] : 0.0;
                ps1 += cm == 1 ? a[q][s] : 0.0;
                ps2 += cm == 2 ? a[q][s] : 0.0;
            }
            ps0 = grp8_sum(ps0);
            ps1 = grp8_sum(ps1);
            ps2 = grp8_sum(ps2);
            if (ti + TI * q < N) rsum = fmax(rsum, fmax(fabs(ps0), fmax(fabs(ps1), fabs(ps2))));
        }
        const double scale = block_max(amax);
        const bool tinv = block_max(rsum) <= translation_tol<T>() * scale * NB && scale > 0.0;
        if (!(scale > 0.0)) {
            // zero matrix: every eigenvalue (0) is below eps -> eps I (clamp), 0 (mirror)
            if (!mirror)
                for (int r = tid; r < N; r += PT) h[ubi(r / 3, r / 3) * 9 + 4 * (r % 3)] = eps;
            if (tid == 0) {
                if (count != nullptr) atomicAdd(count, 1ull);
                if (changed_out != nullptr) changed_out[e] = 1;
            }
            return false;
        }
        if (tinv) {
            const double sh = scale / NB;
#pragma unroll
            for (int q = 0; q < TR; q++)
#pragma unroll
                for (int s = 0; s < TC; s++) {
                    const int r = ti + TI * q, c = tj + TJ * s;
                    if (r < N && c < N && (r % 3) == (c % 3)) a[q][s] += sh;
                }
        }
        // ---- 2. Householder tridiagonalisation in registers --------------------------------
        // a[q][s] at run-time q, s (wave-uniform)
        auto elem_of = [&](int qc, int sc) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < TR; q++)
#pragma unroll
                for (int s = 0; s < TC; s++)
                    if (q == qc && s == sc) v = a[q][s];
            return v;
        };
        // xb / pb entries are read without bounds tests (lane + 64 < 128); entries at N and
        // beyond stay zero, those above the active block are masked by selects
        for (int i = tid; i < 128; i += PT) {
            xb[0][i] = 0.0;
            xb[1][i] = 0.0;
            pb[i] = 0.0;
        }
        __syncthreads();
        if (tj == 0) {
#pragma unroll
            for (int q = 0; q < TR; q++)
                if (ti + TI * q < N) xb[0][ti + TI * q] = a[q][0];
        }
        __syncthreads();
        // Column k of the reduction, for s0 = (k + 1) / TJ fixed at compile time: the column
        // groups s < s0 and the row groups q < q0 = s0 / 4 lie wholly above row k + 1 and drop
        // out of every loop.
        auto step = [&](auto s0c, int k) {
            constexpr int S0 = decltype(s0c)::value, Q0 = S0 * TJ / TI;
            // every LDS read of the step's first half in one batch: column k (rows k + 1 ..
            // N - 1) at the lanes (for the norm), the own rows and columns, and x_{k+1}
            const double* x = xb[k & 1];
            const double xl0 = x[lane], xl1 = x[lane + 64], x0 = x[k + 1];
            double xr[TR], xc[TC];
#pragma unroll
            for (int q = Q0; q < TR; q++) xr[q] = x[ti + TI * q];
#pragma unroll
            for (int s = S0; s < TC; s++) xc[s] = x[tj + TJ * s];
            double ss = (lane >= k + 2 ? xl0 * xl0 : 0.0) + (lane + 64 >= k + 2 ? xl1 * xl1 : 0.0);
            ss = wave_sum_all(ss);
            double beta = x0, t = 0.0, invd = 0.0;
            if (ss != 0.0) {
                beta = -copysign(sqrt(x0 * x0 + ss), x0);
                t = (beta - x0) * rcp_nr(beta);
                invd = rcp_nr(x0 - beta);
            }
            // v: 0 up to k, 1 at k + 1, x invd below
            auto vsel = [&](int i, double xi) { return i == k + 1 ? 1.0 : (i > k + 1 ? xi * invd : 0.0); };
            if (wid == 0) {
                if (lane >= k + 2 && lane < N) V[voff(k) + lane - k - 2] = xl0 * invd;
                if (lane + 64 >= k + 2 && lane + 64 < N) V[voff(k) + lane + 64 - k - 2] = xl1 * invd;
            }
            if (t != 0.0) {
                double vr[TR], vc[TC], p[TR], pc[TC];
#pragma unroll
                for (int q = Q0; q < TR; q++) vr[q] = vsel(ti + TI * q, xr[q]);
#pragma unroll
                for (int s = S0; s < TC; s++) vc[s] = vsel(tj + TJ * s, xc[s]);
                // p = t A22 v for the own rows
#pragma unroll
                for (int q = Q0; q < TR; q++) {
                    double acc = a[q][S0] * vc[S0];
#pragma unroll
                    for (int s = S0 + 1; s < TC; s++) acc += a[q][s] * vc[s];
                    p[q] = t * grp8_sum(acc);
                }
                if (tj == 0) {
#pragma unroll
                    for (int q = Q0; q < TR; q++)
                        if (q < TR - 1 || ti + TI * q < N) pb[ti + TI * q] = p[q];
                }
                __syncthreads();
                const double pl0 = pb[lane], pl1 = pb[lane + 64];
#pragma unroll
                for (int s = S0; s < TC; s++) pc[s] = pb[tj + TJ * s];
                const double vl0 = vsel(lane, xl0), vl1 = vsel(lane + 64, xl1);
                double kk = (lane > k ? pl0 * vl0 : 0.0) + (lane + 64 > k ? pl1 * vl1 : 0.0);
                kk = 0.5 * t * wave_sum_all(kk);
                // w = p - kk v (padded rows and columns, N and beyond, hold zeros throughout)
#pragma unroll
                for (int q = Q0; q < TR; q++) p[q] = ti + TI * q > k ? p[q] - kk * vr[q] : 0.0;
#pragma unroll
                for (int s = S0; s < TC; s++) pc[s] = tj + TJ * s > k ? pc[s] - kk * vc[s] : 0.0;
#pragma unroll
                for (int q = Q0; q < TR; q++)
#pragma unroll
                    for (int s = S0; s < TC; s++) a[q][s] -= vr[q] * pc[s] + p[q] * vc[s];
            }
            // diagonal entry k: column group k / TJ is S0, or S0 - 1 for the phase's first k
            if (ti == (k & (TI - 1)) && tj == (k & (TJ - 1))) {
                if constexpr (S0 > 0) {
                    constexpr int KF = S0 * TJ - 1;  // the phase's first column
                    L.dg[k] = k == KF ? a[KF / TI][KF / TJ] : a[Q0][S0];
                }
                else {
                    L.dg[k] = a[0][0];
                }
            }
            if (tid == 0) {
                L.od[k] = beta;
                L.tau[k] = t;
            }
            // column k + 1 (rows below it; in column group S0, row groups q >= Q0) for the
            // next step; only the last row group runs past N
            if (tj == ((k + 1) & (TJ - 1))) {
#pragma unroll
                for (int q = Q0; q < TR; q++)
                    if (q < TR - 1 || ti + TI * q < N) xb[(k + 1) & 1][ti + TI * q] = a[q][S0];
            }
            __syncthreads();
        };
        static_for<0, TC>([&](auto s0c) {
            constexpr int S0 = decltype(s0c)::value;
            const int kb = S0 == 0 ? 0 : TJ * S0 - 1;
            const int ke = TJ * S0 + TJ - 1 < N - 2 ? TJ * S0 + TJ - 1 : N - 2;
            for (int k = kb; k < ke; k++) step(s0c, k);
        });
        {
            // the last 2 x 2 block: (N-2, N-2), (N-1, N-2), (N-1, N-1)
            const int r0 = N - 2, r1 = N - 1;
            if (ti == (r0 % TI) && tj == (r0 % TJ)) L.dg[r0] = elem_of(r0 / TI, r0 / TJ);
            if (ti == (r1 % TI) && tj == (r0 % TJ)) L.od[r0] = elem_of(r1 / TI, r0 / TJ);
            if (ti == (r1 % TI) && tj == (r1 % TJ)) L.dg[r1] = elem_of(r1 / TI, r1 / TJ);
            if (tid == 0) {
                L.od[N - 1] = 0.0;
                L.tau[N - 2] = 0.0;
                L.tau[N - 1] = 0.0;
            }
        }
        __syncthreads();
        return project_tail<N, false, PT, T, KMAX, PL>(h, e, eps, mirror, count, changed_out, tinv, scale, L,
                                                       [&](int i, int k) { return V[voff(k) + i - k - 2]; });
    }

    template <int N, typename T, int KMAX, int PL>
    __global__ __launch_bounds__(PNT, 3) void k_project_tiled(T* __restrict__ H, int64_t n_mats, int64_t rs, double eps, int mirror,
                                                             unsigned long long* count, const uint8_t* __restrict__ select,
                                                             uint8_t* __restrict__ too_many)
    {
        const int64_t e = blockIdx.x;
        if (e >= n_mats) return;
        bool left = false;
        if (select == nullptr || select[e]) left = project_tiled_one<N, T, KMAX, PL>(H + e * rs, e, eps, mirror, count, nullptr);
        if (too_many != nullptr && threadIdx.x == 0) too_many[e] = left ? 1 : 0;
    }

    // too_many (nullable): 1 for each matrix left to the full-size variant (KMAX < N), 0 for
    // every other one, selected or not.
    template <int N, bool DENSE, typename T, int KMAX = N, int PL = PLU>
    __global__ __launch_bounds__(PNT) void k_project_large(T* __restrict__ H, int64_t n_mats, int64_t rs, double eps, int mirror,
                                                          unsigned long long* count, const uint8_t* __restrict__ select,
                                                          uint8_t* __restrict__ changed_out, uint8_t* __restrict__ too_many)
    {
        const int64_t e = blockIdx.x;
        if (e >= n_mats) return;
        bool left = false;
        if (select == nullptr || select[e])
            left = project_large_one<N, DENSE, PNT, T, KMAX, PL>(H + e * rs, e, eps, mirror, count, changed_out);
        if (too_many != nullptr && threadIdx.x == 0) too_many[e] = left ? 1 : 0;
    }

    // The elements listed in list[0 .. *n_list) (upper-block storage), a fixed grid of
    // PT-thread workgroups striding over the list (one wave per 12x12 matrix).
    template <int N, int PT, typename T>
    __global__ __launch_bounds__(PT) void k_project_large_list(T* __restrict__ H, int64_t rs, const uint32_t* __restrict__ list,
                                                              const unsigned int* __restrict__ n_list, double eps, int mirror,
                                                              unsigned long long* count)
    {
        const unsigned int n = *n_list;
        for (unsigned int i = blockIdx.x; i < n; i += gridDim.x) {
            const int64_t e = list[i];
            project_large_one<N, false, PT, T>(H + e * rs, e, eps, mirror, count, nullptr);
            __syncthreads();  // LDS is reused by the next matrix
        }
    }

    // list[] = indices e with flags[e] != 0, in no particular order; *n_list counts them
    // (zeroed by the caller's k_zero_counter on the same stream).
    __global__ __launch_bounds__(NT) void k_list_flagged(const uint8_t* __restrict__ flags, int64_t n, uint32_t* __restrict__ list,
                                                         unsigned int* __restrict__ n_list)
    {
        const int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (e < n && flags[e]) list[atomicAdd(n_list, 1u)] = (uint32_t)e;
    }
    __global__ void k_zero_counter(unsigned int* c) { *c = 0u; }

    template <typename T>
    int launch_project_ub_t(hipStream_t s, T* H, int64_t n_mats, int32_t nb, int64_t rs, double eps, int32_t mirror, uint8_t* flags,
                            uint32_t* list, unsigned int* n_list, const uint8_t* select, unsigned long long* d_count)
    {
        if (n_mats <= 0) return 0;
        if (rs < (int64_t)nb * (nb + 1) / 2 * 9) return -1;
        const unsigned g = (unsigned)((n_mats + 63) / 64);
        switch (nb) {
        case 1: k_project_ub_full<1, T><<<g, 64, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select); return 0;
        case 2: k_project_ub_full<2, T><<<g, 64, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select); return 0;
        case 3: k_project_ub_full<3, T><<<g, 64, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select); return 0;
        case 4:
            // reduced 9x9 path in registers; the few elements it hands back (more than three
            // eigenvalues below eps, or not translation invariant) are listed and projected by
            // the workgroup-per-matrix kernel
            k_project_ub4_reduced<T><<<g, 64, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, flags, select);
            k_zero_counter<<<1, 1, 0, s>>>(n_list);
            k_list_flagged<<<grid_for(n_mats), NT, 0, s>>>(flags, n_mats, list, n_list);
            k_project_large_list<12, 64, T><<<4096, 64, 0, s>>>(H, rs, list, n_list, eps, mirror, d_count);
            return 0;
        case 8: k_project_large<24, false, T><<<(unsigned)n_mats, PNT, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select, nullptr, nullptr); return 0;
        case 10: k_project_large<30, false, T><<<(unsigned)n_mats, PNT, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select, nullptr, nullptr); return 0;
        case 27:
            // 81x81: the matrix in registers during the tridiagonalisation, eigenvectors for at
            // most 8 eigenvalues below eps and four inverse-iteration lanes keep the LDS at 50 KB,
            // three matrices per CU (the full-size LDS variant takes 155 KB, one per CU); Hex27
            // elements have at most 3 such eigenvalues as a rule, the others are listed and
            // projected by the full-size variant
            k_project_tiled<81, T, 8, 4><<<(unsigned)n_mats, PNT, 0, s>>>(H, n_mats, rs, eps, mirror, d_count, select, flags);
            k_zero_counter<<<1, 1, 0, s>>>(n_list);
            k_list_flagged<<<grid_for(n_mats), NT, 0, s>>>(flags, n_mats, list, n_list);
            k_project_large_list<81, PNT, T><<<1024, PNT, 0, s>>>(H, rs, list, n_list, eps, mirror, d_count);
            return 0;
        default: return -1;
        }
    }

    int launch_project_ub(hipStream_t s, double* H, int64_t n_mats, int32_t nb, int64_t rs, double eps, int32_t mirror, uint8_t* flags,
                          uint32_t* list, unsigned int* n_list, const uint8_t* select, unsigned long long* d_count)
    {
        return launch_project_ub_t(s, H, n_mats, nb, rs, eps, mirror, flags, list, n_list, select, d_count);
    }
    int launch_project_ub(hipStream_t s, float* H, int64_t n_mats, int32_t nb, int64_t rs, double eps, int32_t mirror, uint8_t* flags,
                          uint32_t* list, unsigned int* n_list, const uint8_t* select, unsigned long long* d_count)
    {
        return launch_project_ub_t(s, H, n_mats, nb, rs, eps, mirror, flags, list, n_list, select, d_count);
    }

    // ---- progressive projection (NewtonsMethod.cpp:306-330, ElementHessians.cpp:79-165) ----
    __global__ __launch_bounds__(NT) void k_ppn_select(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                      const int32_t* __restrict__ rows, const double* __restrict__ grad, int32_t nblocks,
                                                      double threshold, PotRanges pr, const uint8_t* __restrict__ projected,
                                                      uint8_t* __restrict__ select, unsigned long long* n_active)
    {
        const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;  // gradient list t, of DoF block rows[t]
        bool active = false;
        if (t < nblocks) {
            const int64_t b = rows[t];
            const double m = fmax(fabs(grad[3 * b]), fmax(fabs(grad[3 * b + 1]), fabs(grad[3 * b + 2])));
            active = m >= threshold;
            if (active) {
                for (uint32_t k = off[t]; k < off[t + 1]; k++) {
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
    void launch_ppn_select(hipStream_t s, const uint32_t* off, const uint32_t* list, const int32_t* rows, const double* grad, int32_t nblocks,
                           double threshold, const PotRanges& pr, const uint8_t* projected, uint8_t* select, unsigned long long* n_active)
    {
        if (nblocks <= 0) return;
        k_ppn_select<<<(unsigned)((nblocks + NT - 1) / NT), NT, 0, s>>>(off, list, rows, grad, nblocks, threshold, pr, projected, select,
                                                                        n_active);
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
    __global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }
    void launch_plane_max_step(hipStream_t s, const int32_t* blocks, int32_t n, const double* x, const double* du, int32_t axis, double plane,
                               double eta, unsigned long long* out)
    {
        if (n <= 0) return;
        // start from the bits of 1.0 (set on the stream: no host round trip)
        k_set_u64<<<1, 1, 0, s>>>(out, 0x3FF0000000000000ull);
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
        case 24: k_project_large<24, true, double><<<(unsigned)n_mats, PNT, 0, s>>>(mats, n_mats, 24 * 24, eps, mirror, nullptr, nullptr, changed, nullptr); return 0;
        case 30: k_project_large<30, true, double><<<(unsigned)n_mats, PNT, 0, s>>>(mats, n_mats, 30 * 30, eps, mirror, nullptr, nullptr, changed, nullptr); return 0;
        case 81: k_project_large<81, true, double><<<(unsigned)n_mats, PNT, 0, s>>>(mats, n_mats, 81 * 81, eps, mirror, nullptr, nullptr, changed, nullptr); return 0;
        default: return -1;
        }
    }

    // ---- assembly: contributions of each block-matrix slot in fixed order --------------
    // List t belongs to slot aslot[t] (slots in the assembly traversal order, whole SELL
    // slices at a time). One lane per block entry: a wave sums 7 lists (63 lanes), so each
    // gathered 3x3 element block is one contiguous 36- or 72-byte read of nine lanes, and a
    // ninth as many lists are in flight as with a thread per list, which keeps the element
    // blocks they share in L2 between the lists that read them. Every entry is summed in
    // list order (the order of the thread-per-list form, bit for bit).
    constexpr int kSlotsPerWave = 7;
    template <typename TS, typename TH>
    __global__ __launch_bounds__(NT) void k_bsr_assemble(int64_t n_slots, const uint32_t* __restrict__ aslot,
                                                         const uint32_t* __restrict__ mirror, const uint32_t* __restrict__ coff,
                                                         const uint32_t* __restrict__ clist, const TH* __restrict__ H,
                                                         TS* __restrict__ val)
    {
        const int lane = threadIdx.x & 63, sub = lane / 9, q = lane - 9 * sub;
        if (sub >= kSlotsPerWave) return;
        const int64_t t = ((xcd_remap() * NT + threadIdx.x) >> 6) * kSlotsPerWave + sub;
        if (t >= n_slots) return;
        const uint32_t slot = aslot[t];
        const int qt = 3 * (q % 3) + q / 3;  // transposed entry
        double acc = 0.0;
        const uint32_t k1 = coff[t + 1];
        uint32_t k = coff[t];
        // kSlotAhead contributions per round: their codes, then their values, load independently
        for (; k + kSlotAhead <= k1; k += kSlotAhead) {
            uint32_t code[kSlotAhead];
            TH v[kSlotAhead];
#pragma unroll
            for (int u = 0; u < kSlotAhead; u++) code[u] = clist[k + u];
#pragma unroll
            for (int u = 0; u < kSlotAhead; u++) v[u] = H[(int64_t)(code[u] >> 1) * 3 + ((code[u] & 1u) ? qt : q)];
#pragma unroll
            for (int u = 0; u < kSlotAhead; u++) acc += (double)v[u];
        }
        for (; k < k1; k++) {
            const uint32_t code = clist[k];
            acc += (double)H[(int64_t)(code >> 1) * 3 + ((code & 1u) ? qt : q)];
        }
        val[sell_val_index<TS>(slot, q)] = (TS)acc;
        // the transposed slot sums the same element blocks transposed, in the same order
        // (checked when the pattern is built), so its sum is this one transposed bit for bit
        if (mirror) {
            const uint32_t m = mirror[t];
            if (m != ~0u) val[sell_val_index<TS>(m, qt)] = (TS)acc;
        }
    }

    void launch_bsr_assemble(hipStream_t s, int64_t n, const uint32_t* aslot, const uint32_t* mirror, const uint32_t* coff,
                             const uint32_t* clist, const double* H, const float* H32, double* v64, float* v32)
    {
        if (n <= 0) return;
        const int64_t waves = (n + kSlotsPerWave - 1) / kSlotsPerWave;
        const unsigned g = (unsigned)((waves + NT / 64 - 1) / (NT / 64));
        if (H32) {
            if (v64) k_bsr_assemble<double, float><<<g, NT, 0, s>>>(n, aslot, mirror, coff, clist, H32, v64);
            if (v32) k_bsr_assemble<float, float><<<g, NT, 0, s>>>(n, aslot, mirror, coff, clist, H32, v32);
            return;
        }
        if (v64) k_bsr_assemble<double, double><<<g, NT, 0, s>>>(n, aslot, mirror, coff, clist, H, v64);
        if (v32) k_bsr_assemble<float, double><<<g, NT, 0, s>>>(n, aslot, mirror, coff, clist, H, v32);
    }

    // Closed-form inverse of the symmetric diagonal block of each row, computed in
    // storage precision like BlockedSparseMatrix.h:1198-1214. The inverse is symmetric by
    // construction, so its six distinct values are stored: dinv[s*n_rows+row] holds entry
    // 0, 1, 2, 4, 5, 8 (s = 0..5) of the row-major 3x3 inverse.
    template <typename TS>
    __global__ __launch_bounds__(NT) void k_dinv(int32_t n_rows, const int64_t* __restrict__ slice_ptr,
                                                 const int32_t* __restrict__ row_at, const TS* __restrict__ val, TS* __restrict__ dinv)
    {
        const int64_t pos = xcd_remap() * NT + threadIdx.x;
        if (pos >= n_rows) return;
        const int64_t row = row_at ? row_at[pos] : pos;
        const int64_t base = slice_ptr[pos >> 6];
        const int64_t t = base + (pos & 63);  // j = 0 is the diagonal block
        const TS m00 = val[sell_val_index<TS>(t, 0)], m01 = val[sell_val_index<TS>(t, 1)], m02 = val[sell_val_index<TS>(t, 2)];
        const TS m11 = val[sell_val_index<TS>(t, 4)], m12 = val[sell_val_index<TS>(t, 5)], m22 = val[sell_val_index<TS>(t, 8)];
        const TS c00 = m11 * m22 - m12 * m12;
        const TS c01 = m02 * m12 - m01 * m22;
        const TS c02 = m01 * m12 - m11 * m02;
        const TS det = m00 * c00 + m01 * c01 + m02 * c02;
        // divided through like BlockedSparseMatrix.h:1198-1214: a singular diagonal block gives
        // inf / NaN entries, and the PCG that uses them ends unconverged, as the reference's does
        TS inv[9];
        const TS id = (TS)(1.0 / (double)det);
        inv[0] = c00 * id;
        inv[1] = c01 * id;
        inv[2] = c02 * id;
        inv[4] = (m00 * m22 - m02 * m02) * id;
        inv[5] = (m01 * m02 - m00 * m12) * id;
        inv[8] = (m00 * m11 - m01 * m01) * id;
        dinv[0 * (int64_t)n_rows + row] = inv[0];
        dinv[1 * (int64_t)n_rows + row] = inv[1];
        dinv[2 * (int64_t)n_rows + row] = inv[2];
        dinv[3 * (int64_t)n_rows + row] = inv[4];
        dinv[4 * (int64_t)n_rows + row] = inv[5];
        dinv[5 * (int64_t)n_rows + row] = inv[8];
    }

    // The full row-major 3x3 block of row `row` from the six stored values.
    template <typename T>
    __device__ __forceinline__ void load_dinv(const T* __restrict__ dinv, int64_t n_rows, int64_t row, T (&d)[9])
    {
        T s[6];
#pragma unroll
        for (int i = 0; i < 6; i++) s[i] = dinv[i * n_rows + row];
        d[0] = s[0], d[1] = s[1], d[2] = s[2];
        d[3] = s[1], d[4] = s[3], d[5] = s[4];
        d[6] = s[2], d[7] = s[4], d[8] = s[5];
    }

    void launch_dinv(hipStream_t s, const SellPattern& P, const double* v64, const float* v32, double* dinv64, float* dinv32)
    {
        if (P.n_rows <= 0) return;
        if (v64 && dinv64) k_dinv<double><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, v64, dinv64);
        if (v32 && dinv32) k_dinv<float><<<grid_for(P.n_rows), NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, v32, dinv32);
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


    template <typename T>
    __device__ __forceinline__ void apply_dinv(const T* __restrict__ dinv, int64_t n_rows, int64_t row, const T (&rv)[3], T (&zv)[3])
    {
        T d[9];
        load_dinv(dinv, n_rows, row, d);
#pragma unroll
        for (int i = 0; i < 3; i++)
            zv[i] = d[3 * i + 0] * rv[0] + d[3 * i + 1] * rv[1] + d[3 * i + 2] * rv[2];
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

    // Start of PCG iteration k: rr, rz of iteration k-1 from the partials, the convergence
    // test of solve_pcg.h:200-210 for iteration k-1 (for k = 1 the early exits of :116-154),
    // beta = rz / rz_prev; workgroup 0 publishes the state. Returns true when the solve is
    // done (every workgroup computes the same values from the same partials).
    template <int NTS>
    __device__ __forceinline__ bool pcg_prologue(int32_t k, PcgState* st, const double* __restrict__ pa_prev,
                                                 const double* __restrict__ pa_prev2, double* lds, double& beta)
    {
        const int64_t G = gridDim.x;
        // rr, rz of iteration k-1 and rz of k-2 (unused for k = 1: pa_prev2 is then scratch)
        double sums[3];
        sum_partials_k<NTS, 3>({ pa_prev, pa_prev + 1, pa_prev2 + 1 }, { 2, 2, 2 }, G, lds, sums);
        const double rr = sums[0];
        const double rz = sums[1];
        beta = 0.0;
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
                beta = rz / sums[2];
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
        return done != 0;
    }

    // 16-byte vectors of the PCG vector passes
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    template <typename T> struct Vec16;
    template <> struct Vec16<float> { using type = f32x4; static constexpr int width = 4; };
    template <> struct Vec16<double> { using type = f64x2; static constexpr int width = 2; };

    // p_new = z + beta p_old over all rows, after the iteration prologue (same grid as the
    // SpMV, whose partials the prologue reads). VEC: 16-byte accesses (the three vectors are
    // 16-byte aligned); the first kDirectionAhead vectors of each thread are loaded before
    // the prologue, whose partial sums would otherwise stall every workgroup first.
    constexpr int kDirectionAhead = 4;
    template <typename TV, int NTS, bool VEC>
    __global__ __launch_bounds__(NTS) void k_pcg_direction(int32_t n_rows, int32_t k, const TV* __restrict__ z,
                                                           const TV* __restrict__ p_old, TV* __restrict__ p_new, PcgState* st,
                                                           const double* __restrict__ pa_prev, const double* __restrict__ pa_prev2)
    {
        __shared__ double lds[3 * NTS / 64];
        if (st->done) return;
        using V = typename Vec16<TV>::type;
        constexpr int W = VEC ? Vec16<TV>::width : 1;
        using L = typename std::conditional<VEC, V, TV>::type;
        const int64_t n = 3 * (int64_t)n_rows;
        const int64_t nv = n / W;
        const int64_t step = (int64_t)gridDim.x * NTS;
        const int64_t i0 = (int64_t)blockIdx.x * NTS + threadIdx.x;
        const L* zv = reinterpret_cast<const L*>(z);
        const L* pv = reinterpret_cast<const L*>(p_old);
        L* out = reinterpret_cast<L*>(p_new);
        L za[kDirectionAhead], pa[kDirectionAhead];
#pragma unroll
        for (int u = 0; u < kDirectionAhead; u++) {
            const int64_t i = i0 + u * step;
            if (i < nv) {
                za[u] = zv[i];
                pa[u] = pv[i];
            }
        }
        double beta;
        if (pcg_prologue<NTS>(k, st, pa_prev, pa_prev2, lds, beta)) return;
        const TV tb = (TV)beta;
        // fma(beta, p_old, z): the contraction of z + beta * p_old elsewhere
#pragma unroll
        for (int u = 0; u < kDirectionAhead; u++) {
            const int64_t i = i0 + u * step;
            if (i < nv) __builtin_nontemporal_store(__builtin_elementwise_fma((L)tb, pa[u], za[u]), out + i);
        }
        for (int64_t i = i0 + kDirectionAhead * step; i < nv; i += step)
            __builtin_nontemporal_store(__builtin_elementwise_fma((L)tb, pv[i], zv[i]), out + i);
        if constexpr (VEC) {
            if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
                const int64_t e = nv * W + threadIdx.x;
                p_new[e] = __builtin_elementwise_fma(tb, p_old[e], z[e]);
            }
        }
    }

    // PRE: the search direction p_new = z + beta p_old was formed by k_pcg_direction, so
    // the gathers read p_new only (long rows, where the gathers of z and p_old cost ~20%).
    template <typename TS, typename TV, bool STREAM, typename CT, int NTS, bool PRE>
    __global__ __launch_bounds__(NTS) void k_pcg_spmv(int32_t n_rows, int32_t k, const int64_t* __restrict__ slice_ptr,
                                                     const int32_t* __restrict__ row_at, const CT* __restrict__ col,
                                                     const TS* __restrict__ val,
                                                     const TV* __restrict__ z, const TV* __restrict__ p_old, TV* __restrict__ p_new,
                                                     TV* __restrict__ q, PcgState* st, const double* __restrict__ pa_prev,
                                                     const double* __restrict__ pa_prev2, double* __restrict__ pb)
    {
        __shared__ double lds[3 * NTS / 64];
        if (st->done) return;
        double beta = 0.0;
        if constexpr (!PRE) {
            if (pcg_prologue<NTS>(k, st, pa_prev, pa_prev2, lds, beta)) return;
        }
        const int64_t G = gridDim.x;
        const TV tb = (TV)beta;
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, G, r0, r1);
        double dot = 0.0;
        for (int64_t pos = r0 + threadIdx.x; pos < r1; pos += NTS) {
            const int64_t row = row_at ? row_at[pos] : pos;
            const int64_t b0 = slice_ptr[pos >> 6];
            const int wd = (int)((slice_ptr[(pos >> 6) + 1] - b0) >> 6);
            const int lane = (int)(pos & 63);
            const CT* cp = col + b0 + lane;
            const TS* vp = val + b0 * 9;
            TV a0 = (TV)0, a1 = (TV)0, a2 = (TV)0;
            // large matrices stream through once per iteration: non-temporal loads keep them
            // from evicting the gathered vector entries (12% faster per PCG iteration on a
            // 6M-tet block; 40% slower on a cache-resident 50k-DoF system, hence STREAM).
            // Unrolled by 2: FP64 with the pass at 80 instead of 113 VGPRs (6 waves per SIMD)
            // made the FP64 PCG iteration 3% (6M-tet block) to 6% (50k-DoF cantilever)
            // faster, FP32 without it 5% faster (200k-DoF ball). FP32 blocks, loaded as two
            // 16-byte runs and one value, take 0.141 ms per iteration on the 6M-tet block
            // unrolled by 2 and 0.154 ms unrolled by 4 (85 VGPRs, 5 waves per SIMD); with
            // nine single-value loads and unrolled by 4 they took 0.143 ms.
            constexpr int kUnroll = 2;
#pragma unroll kUnroll
            for (int j = 0; j < wd; j++) {
                const int64_t c = 3 * block_col(cp[j * 64], row);
                TV x0, x1, x2;
                if constexpr (PRE) {
                    x0 = p_new[c + 0];
                    x1 = p_new[c + 1];
                    x2 = p_new[c + 2];
                }
                else {
                    x0 = z[c + 0] + tb * p_old[c + 0];
                    x1 = z[c + 1] + tb * p_old[c + 1];
                    x2 = z[c + 2] + tb * p_old[c + 2];
                }
                TS v[9];
                load_block<TS, STREAM>(vp + (int64_t)j * 576, lane, v);
                a0 += (TV)v[0] * x0 + (TV)v[1] * x1 + (TV)v[2] * x2;
                a1 += (TV)v[3] * x0 + (TV)v[4] * x1 + (TV)v[5] * x2;
                a2 += (TV)v[6] * x0 + (TV)v[7] * x1 + (TV)v[8] * x2;
            }
            const int64_t kk = 3 * row;
            TV p0, p1, p2;
            if constexpr (PRE) {
                p0 = p_new[kk + 0];
                p1 = p_new[kk + 1];
                p2 = p_new[kk + 2];
            }
            else {
                p0 = z[kk + 0] + tb * p_old[kk + 0];
                p1 = z[kk + 1] + tb * p_old[kk + 1];
                p2 = z[kk + 2] + tb * p_old[kk + 2];
            }
            // non-temporal stores: the vectors go out to memory as they are written instead
            // of staying dirty in the XCD's L2 until the end-of-kernel release, which on the
            // next kernel's critical path writes them back (PCG iteration 0.153 -> 0.148 ms
            // on the 6M-tet block)
            if constexpr (!PRE) {
                __builtin_nontemporal_store(p0, p_new + kk + 0);
                __builtin_nontemporal_store(p1, p_new + kk + 1);
                __builtin_nontemporal_store(p2, p_new + kk + 2);
            }
            __builtin_nontemporal_store(a0, q + kk + 0);
            __builtin_nontemporal_store(a1, q + kk + 1);
            __builtin_nontemporal_store(a2, q + kk + 2);
            dot += (double)p0 * (double)a0 + (double)p1 * (double)a1 + (double)p2 * (double)a2;
        }
        const double s = block_sum<NTS>(dot, lds);
        if (threadIdx.x == 0) pb[w] = s;
    }

    // The same SpMV for small, cache-resident systems, whose workgroups own fewer rows than
    // threads (a 49k-DoF system gives 64 rows to each of 256 workgroups of 512 threads): LPR
    // consecutive lanes share a row, lane g summing the row's blocks j = g, g + LPR, ..., so
    // a row's chain of dependent loads (column id, then the gathered vector) is LPR times
    // shorter; the lanes' partial products are then summed across the LPR lanes in a fixed
    // order (xor butterfly), and the row's first lane stores the result. The host takes this
    // form when every row of a workgroup's chunk gets its LPR lanes in one pass.
    template <typename TS, typename TV, typename CT, int NTS, int LPR, bool PRE>
    __global__ __launch_bounds__(NTS) void k_pcg_spmv_split(int32_t n_rows, int32_t k, const int64_t* __restrict__ slice_ptr,
                                                           const int32_t* __restrict__ row_at, const CT* __restrict__ col,
                                                           const TS* __restrict__ val, const TV* __restrict__ z,
                                                           const TV* __restrict__ p_old, TV* __restrict__ p_new, TV* __restrict__ q,
                                                           PcgState* st, const double* __restrict__ pa_prev,
                                                           const double* __restrict__ pa_prev2, double* __restrict__ pb)
    {
        static_assert(LPR >= 2 && LPR <= 64 && (LPR & (LPR - 1)) == 0, "lanes per row: a power of two");
        // blocks of each lane loaded before the prologue (its partial sums and the state are
        // a round trip of their own; nothing loaded here depends on them)
        constexpr int NPRE = 2;
        __shared__ double lds[3 * NTS / 64];
        const int64_t G = gridDim.x;
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, G, r0, r1);
        const int sub = threadIdx.x & (LPR - 1);
        const int64_t pos = r0 + threadIdx.x / LPR;
        const bool live = pos < r1;  // every lane of a row agrees
        int64_t row = 0, b0 = 0;
        int wd = 0, lane = 0;
        TS v[NPRE][9];
        TV xa[NPRE][3], xb[NPRE][3];
        if (live) {
            row = row_at ? row_at[pos] : pos;
            b0 = slice_ptr[pos >> 6];
            wd = (int)((slice_ptr[(pos >> 6) + 1] - b0) >> 6);
            lane = (int)(pos & 63);
#pragma unroll
            for (int u = 0; u < NPRE; u++) {
                const int j = sub + u * LPR;
                if (j < wd) {
                    const int64_t c = 3 * block_col(col[b0 + lane + (int64_t)j * 64], row);
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if constexpr (PRE) xa[u][i] = p_new[c + i];
                        else {
                            xa[u][i] = z[c + i];
                            xb[u][i] = p_old[c + i];
                        }
                    }
                    load_block<TS, false>(val + b0 * 9 + (int64_t)j * 576, lane, v[u]);
                }
            }
        }
        if (st->done) return;
        double beta = 0.0;
        if constexpr (!PRE) {
            if (pcg_prologue<NTS>(k, st, pa_prev, pa_prev2, lds, beta)) return;
        }
        const TV tb = (TV)beta;
        double dot = 0.0;
        TV a0 = (TV)0, a1 = (TV)0, a2 = (TV)0;
        auto add_block = [&](const TS (&b)[9], TV x0, TV x1, TV x2) {
            a0 += (TV)b[0] * x0 + (TV)b[1] * x1 + (TV)b[2] * x2;
            a1 += (TV)b[3] * x0 + (TV)b[4] * x1 + (TV)b[5] * x2;
            a2 += (TV)b[6] * x0 + (TV)b[7] * x1 + (TV)b[8] * x2;
        };
        if (live) {
#pragma unroll
            for (int u = 0; u < NPRE; u++) {
                if (sub + u * LPR < wd) {
                    if constexpr (PRE) add_block(v[u], xa[u][0], xa[u][1], xa[u][2]);
                    else add_block(v[u], xa[u][0] + tb * xb[u][0], xa[u][1] + tb * xb[u][1], xa[u][2] + tb * xb[u][2]);
                }
            }
            const CT* cp = col + b0 + lane;
            const TS* vp = val + b0 * 9;
#pragma unroll 2
            for (int j = sub + NPRE * LPR; j < wd; j += LPR) {
                const int64_t c = 3 * block_col(cp[j * 64], row);
                TV x0, x1, x2;
                if constexpr (PRE) {
                    x0 = p_new[c + 0];
                    x1 = p_new[c + 1];
                    x2 = p_new[c + 2];
                }
                else {
                    x0 = z[c + 0] + tb * p_old[c + 0];
                    x1 = z[c + 1] + tb * p_old[c + 1];
                    x2 = z[c + 2] + tb * p_old[c + 2];
                }
                TS b[9];
                load_block<TS, false>(vp + (int64_t)j * 576, lane, b);
                add_block(b, x0, x1, x2);
            }
        }
        // every lane takes part in the butterfly (rows of a wave are whole: LPR divides 64)
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1) {
            a0 += __shfl_xor(a0, m, 64);
            a1 += __shfl_xor(a1, m, 64);
            a2 += __shfl_xor(a2, m, 64);
        }
        if (live && sub == 0) {
            const int64_t kk = 3 * row;
            TV p0, p1, p2;
            if constexpr (PRE) {
                p0 = p_new[kk + 0];
                p1 = p_new[kk + 1];
                p2 = p_new[kk + 2];
            }
            else {
                p0 = z[kk + 0] + tb * p_old[kk + 0];
                p1 = z[kk + 1] + tb * p_old[kk + 1];
                p2 = z[kk + 2] + tb * p_old[kk + 2];
                p_new[kk + 0] = p0;
                p_new[kk + 1] = p1;
                p_new[kk + 2] = p2;
            }
            q[kk + 0] = a0;
            q[kk + 1] = a1;
            q[kk + 2] = a2;
            dot = (double)p0 * (double)a0 + (double)p1 * (double)a1 + (double)p2 * (double)a2;
        }
        const double s = block_sum<NTS>(dot, lds);
        if (threadIdx.x == 0) pb[w] = s;
    }

    // Rows loaded by each thread of k_pcg_update before its partial sums (two rows per
    // thread cover a 1024-row chunk of a 512-thread workgroup).
    constexpr int kUpdateAhead = 2;
    template <typename TV, int NTS>
    __global__ __launch_bounds__(NTS) void k_pcg_update(int32_t n_rows, int32_t k, const TV* __restrict__ dinv, TV* __restrict__ x,
                                                       TV* __restrict__ r, TV* __restrict__ z, const TV* __restrict__ p,
                                                       const TV* __restrict__ q, PcgState* st, const double* __restrict__ pa_cur,
                                                       double* __restrict__ pa_next, const double* __restrict__ pb)
    {
        __shared__ double lds[2 * NTS / 64];
        const int64_t G = gridDim.x;
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, G, r0, r1);
        TV xa[kUpdateAhead][3], ra[kUpdateAhead][3], pa[kUpdateAhead][3], qa[kUpdateAhead][3], da[kUpdateAhead][9];
#pragma unroll
        for (int u = 0; u < kUpdateAhead; u++) {
            const int64_t row = r0 + threadIdx.x + u * NTS;
            if (row < r1) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    xa[u][i] = x[3 * row + i];
                    ra[u][i] = r[3 * row + i];
                    pa[u][i] = p[3 * row + i];
                    qa[u][i] = q[3 * row + i];
                }
                load_dinv(dinv, n_rows, row, da[u]);
            }
        }
        // the state is read after the rows are requested (one round trip less; the loads
        // read valid memory whether or not the solve is done)
        if (st->done) return;
        double sums[2];
        sum_partials_k<NTS, 2>({ pb, pa_cur + 1 }, { 1, 2 }, G, lds, sums);
        const double pAp = sums[0];
        const double rz = sums[1];
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
        double rr = 0.0, rzn = 0.0;
        // x += alpha p, r -= alpha q, z = D^-1 r (non-temporal stores as in k_pcg_spmv)
        auto finish = [&](int64_t row, const TV (&xv)[3], const TV (&rv0)[3], const TV (&pv)[3], const TV (&qv)[3], const TV (&d)[9]) {
            const int64_t kk = 3 * row;
            TV rv[3], zv[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                __builtin_nontemporal_store(xv[i] + ta * pv[i], x + kk + i);
                rv[i] = rv0[i] - ta * qv[i];
                __builtin_nontemporal_store(rv[i], r + kk + i);
            }
#pragma unroll
            for (int i = 0; i < 3; i++) zv[i] = d[3 * i + 0] * rv[0] + d[3 * i + 1] * rv[1] + d[3 * i + 2] * rv[2];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                __builtin_nontemporal_store(zv[i], z + kk + i);
                rr += (double)rv[i] * (double)rv[i];
                rzn += (double)rv[i] * (double)zv[i];
            }
        };
#pragma unroll
        for (int u = 0; u < kUpdateAhead; u++) {
            const int64_t row = r0 + threadIdx.x + u * NTS;
            if (row < r1) finish(row, xa[u], ra[u], pa[u], qa[u], da[u]);
        }
        for (int64_t row = r0 + threadIdx.x + kUpdateAhead * NTS; row < r1; row += NTS) {
            TV xv[3], rv[3], pv[3], qv[3], d[9];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                xv[i] = x[3 * row + i];
                rv[i] = r[3 * row + i];
                pv[i] = p[3 * row + i];
                qv[i] = q[3 * row + i];
            }
            load_dinv(dinv, n_rows, row, d);
            finish(row, xv, rv, pv, qv, d);
        }
        double red[2] = { rr, rzn };
        block_sum_k<NTS, 2>(red, lds);
        if (threadIdx.x == 0) {
            pa_next[2 * w + 0] = red[0];
            pa_next[2 * w + 1] = red[1];
        }
    }

    // PCG grid: at most one full wave of workgroups at the SpMV's occupancy (3 workgroups of
    // 512 threads per CU at 80 VGPRs, 256 CUs); with 1024 the last 256 ran as a tail. Hex27
    // 40^3 0.316 -> 0.280 ms per iteration (FP64 0.62 -> 0.57), 6M-tet block 1-2% faster.
    constexpr int kPcgMaxGroups = 768;
    // Small systems get at least one workgroup per CU (256) while that leaves 64 rows per
    // workgroup: the 49k-DoF cantilever (16k rows, 64 workgroups by the reduction rule) runs
    // a PCG iteration in 13.5 instead of 16.5 us; the 214k-DoF ball (279 workgroups) is
    // slower with 768 (23 instead of 20 us), so the rule only raises the count up to 256.
    int pcg_grid(int32_t n_rows)
    {
        int g = reduction_groups(n_rows);
        const int per_cu = (n_rows + 63) / 64;
        if (g < 256) g = per_cu < 256 ? (per_cu > g ? per_cu : g) : 256;
        return g > kPcgMaxGroups ? kPcgMaxGroups : g;
    }

    // Instantiation by matrix size (streaming loads), column-id width and direction pass.
    template <typename TS, typename TV, bool STREAM, typename CT>
    void launch_spmv_c(hipStream_t s, const SellPattern& P, const CT* col, bool pre, int G, int32_t k, const TS* val, const TV* z,
                       const TV* p_old, TV* p_new, TV* q, PcgState* st, const double* pa_prev, const double* pa_prev2, double* pb)
    {
        if (pre)
            k_pcg_spmv<TS, TV, STREAM, CT, kSpmvThreads, true><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_row_at, col, val, z,
                                                                                         p_old, p_new, q, st, pa_prev, pa_prev2, pb);
        else
            k_pcg_spmv<TS, TV, STREAM, CT, kSpmvThreads, false><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_row_at, col, val, z,
                                                                                          p_old, p_new, q, st, pa_prev, pa_prev2, pb);
    }
    // Lanes per row of the split SpMV for a cache-resident system: the rows of a workgroup's
    // chunk (row_chunk) times LPR must fit its threads in one pass; 1: the one-lane form.
    inline int spmv_lanes_per_row(int32_t n_rows, int G)
    {
        const int64_t per = ((((int64_t)n_rows + G - 1) / G) + 63) / 64 * 64;
        int lpr = 1;
        while (lpr < 8 && per * (2 * lpr) <= kSpmvThreads) lpr *= 2;
        return lpr;
    }
    template <typename TS, typename TV, typename CT, int LPR>
    void launch_spmv_split(hipStream_t s, const SellPattern& P, const CT* col, bool pre, int G, int32_t k, const TS* val, const TV* z,
                           const TV* p_old, TV* p_new, TV* q, PcgState* st, const double* pa_prev, const double* pa_prev2, double* pb)
    {
        if (pre)
            k_pcg_spmv_split<TS, TV, CT, kSpmvThreads, LPR, true><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_row_at, col, val, z,
                                                                                            p_old, p_new, q, st, pa_prev, pa_prev2, pb);
        else
            k_pcg_spmv_split<TS, TV, CT, kSpmvThreads, LPR, false><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, P.d_slice_ptr, P.d_row_at, col, val, z,
                                                                                             p_old, p_new, q, st, pa_prev, pa_prev2, pb);
    }
    template <typename TS, typename TV, typename CT>
    void launch_spmv_small(hipStream_t s, const SellPattern& P, const CT* col, bool pre, int G, int lpr, int32_t k, const TS* val,
                           const TV* z, const TV* p_old, TV* p_new, TV* q, PcgState* st, const double* pa_prev, const double* pa_prev2,
                           double* pb)
    {
        switch (lpr) {
        case 2: launch_spmv_split<TS, TV, CT, 2>(s, P, col, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb); break;
        case 4: launch_spmv_split<TS, TV, CT, 4>(s, P, col, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb); break;
        default: launch_spmv_split<TS, TV, CT, 8>(s, P, col, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb); break;
        }
    }
    template <typename TS, typename TV>
    void launch_spmv(hipStream_t s, const SellPattern& P, int G, int32_t k, const TS* val, const TV* z, const TV* p_old, TV* p_new,
                     TV* q, PcgState* st, const double* pa_prev, const double* pa_prev2, double* pb)
    {
        // Long rows (27-node hexahedra) or a streamed matrix: form the direction first so the
        // SpMV gathers one vector (Hex27 -10% FP32 / -30% FP64, 6M-tet block -1% / -8.5%). On
        // cache-resident systems the extra launch costs more than the gathers it saves.
        const bool stream = (double)P.n_slots * 9 * sizeof(TS) > kStreamBytes;
        const bool pre = stream || P.nnzb >= (int64_t)kDirectionPassMeanRow * P.n_rows;
        if (pre) {
            if ((((uintptr_t)z | (uintptr_t)p_old | (uintptr_t)p_new) & 15) == 0)
                k_pcg_direction<TV, kSpmvThreads, true><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, z, p_old, p_new, st, pa_prev, pa_prev2);
            else
                k_pcg_direction<TV, kSpmvThreads, false><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, z, p_old, p_new, st, pa_prev, pa_prev2);
        }
        const int lpr = stream ? 1 : spmv_lanes_per_row(P.n_rows, G);
        if (lpr > 1) {
            if (P.d_dcol) launch_spmv_small<TS, TV>(s, P, P.d_dcol, pre, G, lpr, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
            else launch_spmv_small<TS, TV>(s, P, P.d_col, pre, G, lpr, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
        }
        else if (P.d_dcol) {
            if (stream) launch_spmv_c<TS, TV, true>(s, P, P.d_dcol, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
            else launch_spmv_c<TS, TV, false>(s, P, P.d_dcol, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
        }
        else {
            if (stream) launch_spmv_c<TS, TV, true>(s, P, P.d_col, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
            else launch_spmv_c<TS, TV, false>(s, P, P.d_col, pre, G, k, val, z, p_old, p_new, q, st, pa_prev, pa_prev2, pb);
        }
    }

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
        launch_spmv<double, double>(s, P, G, k, val, z, p_old, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
        k_pcg_update<double, kSpmvThreads><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, dinv, x, r, z, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
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
        launch_spmv<float, float>(s, P, G, k, val, z, p_old, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
        k_pcg_update<float, kSpmvThreads><<<G, kSpmvThreads, 0, s>>>(P.n_rows, k, dinv, x, r, z, p_new, q, st, pa[(k - 1) & 1], pa[k & 1], pb);
    }

    // ---- mixed-precision refinement -----------------------------------------------
    // r = b - A x with FP32 matrix values, FP64 vectors and FP64 accumulation (the
    // outer residual of iterative refinement); r32 = (float) r; *rr_out = ||r||^2.
    template <bool STREAM, typename CT>
    __global__ __launch_bounds__(NT) void k_residual_mixed(int32_t n_rows, const int64_t* __restrict__ slice_ptr,
                                                           const int32_t* __restrict__ row_at, const CT* __restrict__ col,
                                                           const float* __restrict__ val,
                                                           const double* __restrict__ b, const double* __restrict__ x,
                                                           double* __restrict__ r, float* __restrict__ r32, double* __restrict__ partials)
    {
        __shared__ double lds[NT / 64];
        const int64_t w = xcd_remap();
        int64_t r0, r1;
        row_chunk(n_rows, w, gridDim.x, r0, r1);
        double rr = 0.0;
        for (int64_t pos = r0 + threadIdx.x; pos < r1; pos += NT) {
            const int64_t row = row_at ? row_at[pos] : pos;
            const int64_t b0 = slice_ptr[pos >> 6];
            const int wd = (int)((slice_ptr[(pos >> 6) + 1] - b0) >> 6);
            const int lane = (int)(pos & 63);
            const CT* cp = col + b0 + lane;
            const float* vp = val + b0 * 9;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 4
            for (int j = 0; j < wd; j++) {
                const int64_t c = 3 * block_col(cp[j * 64], row);
                const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
                float v[9];
                load_block<float, STREAM>(vp + (int64_t)j * 576, lane, v);
                a0 += (double)v[0] * x0 + (double)v[1] * x1 + (double)v[2] * x2;
                a1 += (double)v[3] * x0 + (double)v[4] * x1 + (double)v[5] * x2;
                a2 += (double)v[6] * x0 + (double)v[7] * x1 + (double)v[8] * x2;
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
        const bool stream = (double)P.n_slots * 9 * sizeof(float) > kStreamBytes;
        if (P.d_dcol) {
            if (stream) k_residual_mixed<true, int16_t><<<G, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, P.d_dcol, val, b, x, r, r32, partials);
            else k_residual_mixed<false, int16_t><<<G, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, P.d_dcol, val, b, x, r, r32, partials);
        }
        else {
            if (stream) k_residual_mixed<true, int32_t><<<G, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, P.d_col, val, b, x, r, r32, partials);
            else k_residual_mixed<false, int32_t><<<G, NT, 0, s>>>(P.n_rows, P.d_slice_ptr, P.d_row_at, P.d_col, val, b, x, r, r32, partials);
        }
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

// ---- allocation-free loop evaluation (symx_hip_loop_eval_noalloc) ----------------------
// The transposed connectivity [column][element] that the generated loop kernels gather
// through. One copy per GPU for the life of the process; every call on a GPU shares it.
static __device__ __align__(256) int32_t scratch[SYMX_HIP_LOOP_SCRATCH_ENTRIES];

namespace symx_hip
{
    __global__ __launch_bounds__(NT) void k_transpose_conn(const int32_t* __restrict__ in, int64_t n_elem, int32_t stride,
                                                           int32_t* __restrict__ out)
    {
        const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (t >= n_elem * stride) return;
        const int64_t e = t / stride, c = t - e * stride;
        out[c * n_elem + e] = in[t];
    }
    void launch_transpose_conn(hipStream_t s, const int32_t* in, int32_t n_elem, int32_t stride, int32_t* out)
    {
        const int64_t n = (int64_t)n_elem * stride;
        if (n > 0) k_transpose_conn<<<grid_for(n), NT, 0, s>>>(in, n_elem, stride, out);
    }

    // Device address of `scratch` on `device`, which must be the current device. Looked up
    // once per device (each GPU holds its own copy at its own address); the host-side name
    // of the array is no device address and must not reach a kernel or a copy.
    int loop_scratch_address(int device, int32_t** out)
    {
        static std::mutex m;
        static int32_t* table[64] = {};
        std::lock_guard<std::mutex> g(m);
        if (device < 0 || device >= 64) return -1;
        if (table[device] == nullptr) {
            void* ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) return -1;
            table[device] = static_cast<int32_t*>(ptr);
        }
        *out = table[device];
        return 0;
    }
}
