// Batched dense complex LU with partial pivoting for libhbem (gfx950).
//
// Replaces the reference's CPU Accelerate cgesv/cgetrs/cgecon call sites
// (main.swift:7834-7941, 475-503, 6964-7070) with a right-looking blocked
// LU on the GPU.  Every kernel handles a batch of independent systems
// (grid.y / grid.z = system) so one launch serves all resident frequencies.
//
// Layout: column-major interleaved complex float, element (r, c) of system b
// at float offset 2 * (b * a_stride + c * lda + r).  Matrices are padded to
// order N (a multiple of kLuPad) with an identity block, so no kernel needs
// edge handling; padded rows are zero in every true column and never win a
// pivot search of a non-singular matrix.
//
// Pivoting follows LAPACK cgetf2: the pivot of a column is the first row of
// maximal |re| + |im| (icamax), ipiv[j] is the row swapped with row j, and
// swaps are applied to the whole row (left factor columns included).
#include <hip/hip_runtime.h>

#include "hbem_internal.h"

namespace hbem {
namespace {

constexpr int kNB = 32;  // panel width

struct c32 {
    float re, im;
};

__device__ __forceinline__ c32 cmul(c32 a, c32 b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ c32 cconj(c32 a) { return {a.re, -a.im}; }
__device__ __forceinline__ c32 cinv(c32 a) {
    const float d = a.re * a.re + a.im * a.im;
    return {a.re / d, -a.im / d};
}

__device__ __forceinline__ c32 ld(const float *A, size_t idx) {
    const float2 v = *reinterpret_cast<const float2 *>(A + 2 * idx);
    return {v.x, v.y};
}
__device__ __forceinline__ void st(float *A, size_t idx, c32 v) {
    *reinterpret_cast<float2 *>(A + 2 * idx) = make_float2(v.re, v.im);
}

// ---------------------------------------------------------------------------
// Padding and norms.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_pad_identity(float *A, int64_t lda, int64_t a_stride,
                                                      int n, int N) {
    const int c = blockIdx.x;
    const int b = blockIdx.y;
    float *Ab = A + (size_t)b * a_stride * 2;
    for (int r = threadIdx.x; r < N; r += 256) {
        if (r < n && c < n) continue;
        st(Ab, (size_t)c * lda + r, c32{r == c ? 1.0f : 0.0f, 0.0f});
    }
}

// max_c sum_r |a_rc| over the true n x n block (clange '1').  One wave per
// column; norms[] must be zero on entry.
__global__ __launch_bounds__(256) void k_norm1(const float *A, int64_t lda, int64_t a_stride, int n,
                                               double *norms) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int b = blockIdx.y;
    if (c >= n) return;
    const float *Ab = A + (size_t)b * a_stride * 2;
    double s = 0.0;
    for (int r = lane; r < n; r += 64) {
        const c32 v = ld(Ab, (size_t)c * lda + r);
        s += sqrt((double)v.re * v.re + (double)v.im * v.im);
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(norms + b),
                  static_cast<unsigned long long>(__double_as_longlong(s)));
}

// ---------------------------------------------------------------------------
// Panel factorisation: one 1024-thread block per system factors columns
// [k, k + nb) over rows [k, N).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_lu_panel(float *A, int64_t lda, int64_t a_stride, int N,
                                                   int k, int nb, int32_t *ipiv, int32_t *info) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    float *Ab = A + (size_t)b * a_stride * 2;
    __shared__ float red_v[16];
    __shared__ int red_i[16];
    __shared__ int piv_row;
    __shared__ c32 urow[kNB];
    for (int j = 0; j < nb; ++j) {
        const int col = k + j;
        float best = -1.0f;
        int bi = N;
        for (int r = col + tid; r < N; r += 1024) {
            const c32 v = ld(Ab, (size_t)col * lda + r);
            const float a = fabsf(v.re) + fabsf(v.im);
            if (a > best) {
                best = a;
                bi = r;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = best;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            float bv = red_v[0];
            int bidx = red_i[0];
            for (int w = 1; w < 16; ++w)
                if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bidx)) {
                    bv = red_v[w];
                    bidx = red_i[w];
                }
            if (bidx >= N) bidx = col;
            piv_row = bidx;
            ipiv[(size_t)b * N + col] = bidx;
            if (!(bv > 0.0f) && info[b] == 0) info[b] = col + 1;
        }
        __syncthreads();
        const int p = piv_row;
        if (p != col && tid < nb) {
            const size_t cc = (size_t)(k + tid) * lda;
            const c32 x = ld(Ab, cc + col), y = ld(Ab, cc + p);
            st(Ab, cc + col, y);
            st(Ab, cc + p, x);
        }
        __syncthreads();
        const c32 piv = ld(Ab, (size_t)col * lda + col);
        if (tid < nb - j - 1) urow[tid] = ld(Ab, (size_t)(col + 1 + tid) * lda + col);
        __syncthreads();
        const bool nz = (piv.re != 0.0f || piv.im != 0.0f);
        const c32 inv = nz ? cinv(piv) : c32{0.0f, 0.0f};
        const int rest = nb - j - 1;
        for (int r = col + 1 + tid; r < N; r += 1024) {
            c32 l = ld(Ab, (size_t)col * lda + r);
            if (nz) {
                l = cmul(l, inv);
                st(Ab, (size_t)col * lda + r, l);
            }
            for (int c = 0; c < rest; ++c) {
                const size_t idx = (size_t)(col + 1 + c) * lda + r;
                const c32 u = urow[c];
                c32 a = ld(Ab, idx);
                a.re -= l.re * u.re - l.im * u.im;
                a.im -= l.re * u.im + l.im * u.re;
                st(Ab, idx, a);
            }
        }
        __syncthreads();
    }
}

// Register-resident panel for N <= 1024: thread t owns row k + t of the
// nb = 32 panel columns; two block barriers per column (reduction, row
// exchange through LDS).  Same pivot rule as k_lu_panel.
__global__ __launch_bounds__(1024) void k_lu_panel_reg(float *A, int64_t lda, int64_t a_stride,
                                                       int N, int k, int32_t *ipiv,
                                                       int32_t *info) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int row = k + tid;
    const bool own = row < N;
    c32 v[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        v[c] = own ? ld(Ab, (size_t)(k + c) * lda + row) : c32{0.0f, 0.0f};
    __shared__ float red_v[16];
    __shared__ int red_i[16];
    __shared__ c32 prow[2][kNB];
    __shared__ c32 orow[2][kNB];
    for (int j = 0; j < kNB; ++j) {
        const int col = k + j;
        // v[j] by predicated selection keeps every register index static
        c32 vj = v[0];
#pragma unroll
        for (int c = 1; c < kNB; ++c)
            if (c == j) vj = v[c];
        float best = -1.0f;
        int bi = N;
        if (own && row >= col) {
            best = fabsf(vj.re) + fabsf(vj.im);
            bi = row;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = best;
            red_i[wave] = bi;
        }
        __syncthreads();
        float bv = red_v[0];
        int p = red_i[0];
#pragma unroll
        for (int w = 1; w < 16; ++w)
            if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < p)) {
                bv = red_v[w];
                p = red_i[w];
            }
        if (p >= N) p = col;
        const int buf = j & 1;
        if (row == p) {
#pragma unroll
            for (int c = 0; c < kNB; ++c) prow[buf][c] = v[c];
        }
        if (row == col && p != col) {
#pragma unroll
            for (int c = 0; c < kNB; ++c) orow[buf][c] = v[c];
        }
        if (tid == 0) {
            ipiv[(size_t)b * N + col] = p;
            if (!(bv > 0.0f) && info[b] == 0) info[b] = col + 1;
        }
        __syncthreads();
        if (row == col) {
#pragma unroll
            for (int c = 0; c < kNB; ++c) v[c] = prow[buf][c];
        } else if (row == p) {
#pragma unroll
            for (int c = 0; c < kNB; ++c) v[c] = orow[buf][c];
        }
        const c32 piv = prow[buf][j];
        const bool nz = (piv.re != 0.0f || piv.im != 0.0f);
        if (own && row > col) {
            c32 l = v[0];
#pragma unroll
            for (int c = 1; c < kNB; ++c)
                if (c == j) l = v[c];
            if (nz) l = cmul(l, cinv(piv));
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                if (c == j) v[c] = l;
                if (c > j) {
                    const c32 u = prow[buf][c];
                    v[c].re -= l.re * u.re - l.im * u.im;
                    v[c].im -= l.re * u.im + l.im * u.re;
                }
            }
        }
    }
    if (own) {
#pragma unroll
        for (int c = 0; c < kNB; ++c) st(Ab, (size_t)(k + c) * lda + row, v[c]);
    }
}

// Row swaps of panel [k, k + nb) on every column outside the panel, then
// the unit-lower solve U12 = L11^{-1} A12 on columns >= k + nb.  The nb
// sequential swaps are composed once per block into a gather list (<= 2 nb
// rows), so each column issues all its loads before any store.
__global__ __launch_bounds__(256) void k_lu_swap_trsm(float *A, int64_t lda, int64_t a_stride,
                                                      int N, int k, int nb,
                                                      const int32_t *ipiv) {
    const int b = blockIdx.y;
    float *Ab = A + (size_t)b * a_stride * 2;
    __shared__ c32 L11[kNB][kNB + 1];
    __shared__ int trow[2 * kNB];  // destination rows (first nb: k..k+nb-1)
    __shared__ int tsrc[2 * kNB];  // source rows
    __shared__ int tcount;
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e % nb, j = e / nb;
        L11[i][j] = ld(Ab, (size_t)(k + j) * lda + k + i);
    }
    if (threadIdx.x == 0) {
        int cnt = nb;
        for (int j = 0; j < nb; ++j) {
            trow[j] = k + j;
            tsrc[j] = k + j;
        }
        for (int j = 0; j < nb; ++j) {
            const int p = ipiv[(size_t)b * N + k + j];
            if (p == k + j) continue;
            int ip = -1;
            for (int m = 0; m < cnt; ++m)
                if (trow[m] == p) ip = m;
            if (ip < 0) {
                ip = cnt++;
                trow[ip] = p;
                tsrc[ip] = p;
            }
            const int t = tsrc[j];  // swap the sources of positions k+j and p
            tsrc[j] = tsrc[ip];
            tsrc[ip] = t;
        }
        tcount = cnt;
    }
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= N || (c >= k && c < k + nb)) return;
    const size_t cc = (size_t)c * lda;
    const int cnt = tcount;
    c32 vals[2 * kNB];
#pragma unroll
    for (int m = 0; m < 2 * kNB; ++m)
        if (m < cnt) vals[m] = ld(Ab, cc + tsrc[m]);
    if (c >= k + nb) {
        // forward substitution with unit L11 on the first nb (= kNB) rows
#pragma unroll
        for (int i = 1; i < kNB; ++i) {
            c32 acc = vals[i];
#pragma unroll
            for (int j = 0; j < i; ++j) {
                const c32 l = L11[i][j];
                acc.re -= l.re * vals[j].re - l.im * vals[j].im;
                acc.im -= l.re * vals[j].im + l.im * vals[j].re;
            }
            vals[i] = acc;
        }
    }
#pragma unroll
    for (int m = 0; m < 2 * kNB; ++m)
        if (m < cnt) st(Ab, cc + trow[m], vals[m]);
}

// Trailing update A22 -= L21 * U12 over rows/cols [k + nb, N), 64 x 64
// output tiles, 256 threads with a 4 x 4 complex micro-tile each.
__global__ __launch_bounds__(256) void k_lu_gemm(float *A, int64_t lda, int64_t a_stride, int N,
                                                 int k, int nb) {
    const int b = blockIdx.z;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int s = k + nb;
    const int r0 = s + blockIdx.x * 64;
    const int c0 = s + blockIdx.y * 64;
    __shared__ c32 Ls[kNB][64];
    __shared__ c32 Us[kNB][64 + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * 64; e += 256) {
        const int rr = e % 64, l = e / 64;
        Ls[l][rr] = (r0 + rr < N) ? ld(Ab, (size_t)(k + l) * lda + r0 + rr) : c32{0.0f, 0.0f};
    }
    for (int e = tid; e < nb * 64; e += 256) {
        const int l = e % nb, cc = e / nb;
        Us[l][cc] = (c0 + cc < N) ? ld(Ab, (size_t)(c0 + cc) * lda + k + l) : c32{0.0f, 0.0f};
    }
    __syncthreads();
    const int tr = tid % 16, tc = tid / 16;
    c32 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = c32{0.0f, 0.0f};
    for (int l = 0; l < nb; ++l) {
        c32 a[4], u[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Ls[l][tr + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = Us[l][tc + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j].re = fmaf(a[i].re, u[j].re, acc[i][j].re);
                acc[i][j].re = fmaf(-a[i].im, u[j].im, acc[i][j].re);
                acc[i][j].im = fmaf(a[i].re, u[j].im, acc[i][j].im);
                acc[i][j].im = fmaf(a[i].im, u[j].re, acc[i][j].im);
            }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (c0 + tc + 16 * j >= N) continue;
        const size_t cc = (size_t)(c0 + tc + 16 * j) * lda;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (r0 + tr + 16 * i >= N) continue;
            const size_t idx = cc + r0 + tr + 16 * i;
            c32 v = ld(Ab, idx);
            v.re -= acc[i][j].re;
            v.im -= acc[i][j].im;
            st(Ab, idx, v);
        }
    }
}

// MFMA trailing update A22 -= L21 * U12 (FP32 matrix cores,
// v_mfma_f32_16x16x4_f32).  64 x 64 complex output tile per block; each of
// the 4 waves owns a 32 x 32 sub-tile as 2 x 2 MFMA tiles with separate real
// and imaginary accumulators (re += Lr Ur - Li Ui, im += Lr Ui + Li Ur).
// The MFMA computes the transposed tile C^T = U^T L^T so that the 16 lanes
// of an output row group cover 16 consecutive matrix rows (coalesced
// column-major stores).  Operands are staged in LDS as separate re/im planes
// with an 80-float row stride (conflict-free reads of the 4 k-rows a wave
// touches per instruction).
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kGS = 80;

__global__ __launch_bounds__(256) void k_lu_gemm_mfma(float *A, int64_t lda, int64_t a_stride,
                                                      int N, int k) {
    const int b = blockIdx.z;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int s = k + kNB;
    const int r0 = s + blockIdx.x * 64;
    const int c0 = s + blockIdx.y * 64;
    __shared__ float Lr[kNB][kGS], Li[kNB][kGS], Ur[kNB][kGS], Ui[kNB][kGS];
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < (kNB * 64) / 256; ++it) {
        const int e = tid + 256 * it;
        const int kk = e / 64, rr = e % 64;
        const c32 v = (r0 + rr < N) ? ld(Ab, (size_t)(k + kk) * lda + r0 + rr) : c32{0.f, 0.f};
        Lr[kk][rr] = v.re;
        Li[kk][rr] = v.im;
    }
#pragma unroll
    for (int it = 0; it < (kNB * 64) / 256; ++it) {
        const int e = tid + 256 * it;
        const int cc = e / kNB, kk = e % kNB;
        const c32 v = (c0 + cc < N) ? ld(Ab, (size_t)(c0 + cc) * lda + k + kk) : c32{0.f, 0.f};
        Ur[kk][cc] = v.re;
        Ui[kk][cc] = v.im;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    f32x4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            accr[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            acci[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
    for (int ks = 0; ks < kNB; ks += 4) {
        const int kk = ks + lk;
        float ur[2], ui[2], lr[2], lim[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            ur[t] = Ur[kk][wc + 16 * t + li];
            ui[t] = Ui[kk][wc + 16 * t + li];
            lr[t] = Lr[kk][wr + 16 * t + li];
            lim[t] = Li[kk][wr + 16 * t + li];
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                // D^T[col][row]: A operand = U^T (col, k), B operand = L^T (k, row)
                accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[ti], lr[tj], accr[ti][tj], 0, 0, 0);
                accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui[ti], lim[tj], accr[ti][tj], 0, 0, 0);
                acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[ti], lr[tj], acci[ti][tj], 0, 0, 0);
                acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[ti], lim[tj], acci[ti][tj], 0, 0, 0);
            }
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = c0 + wc + 16 * ti + 4 * lk + v;
                const int row = r0 + wr + 16 * tj + li;
                if (row < N && col < N) {
                    const size_t idx = (size_t)col * lda + row;
                    c32 x = ld(Ab, idx);
                    x.re -= accr[ti][tj][v];
                    x.im -= acci[ti][tj][v];
                    st(Ab, idx, x);
                }
            }
}

// ---------------------------------------------------------------------------
// Triangular solves, one 1024-thread block per system, vector in LDS.
// ---------------------------------------------------------------------------

__device__ __forceinline__ c32 block_sum(c32 v, c32 *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o >= 1; o >>= 1) {
        v.re += __shfl_xor(v.re, o);
        v.im += __shfl_xor(v.im, o);
    }
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    c32 s{0.0f, 0.0f};
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) {
        s.re += red[w].re;
        s.im += red[w].im;
    }
    return s;
}

// x <- A^{-1} x with the LU factors: pivots, then L and U sweeps in blocks
// of 32 rows.  Wave 0 solves each 32 x 32 diagonal block with shuffles; all
// threads update the remaining rows.  One block per system, x in LDS.
__global__ __launch_bounds__(1024) void k_lu_solve_blocked(const float *A, int64_t lda,
                                                           int64_t a_stride, int N,
                                                           const int32_t *ipiv, float *X,
                                                           int64_t ldx) {
    extern __shared__ c32 xs[];
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int32_t *pv = ipiv + (size_t)b * N;
    float *xb = X + (size_t)b * ldx * 2;
    for (int r = tid; r < N; r += 1024) xs[r] = ld(xb, r);
    __syncthreads();
    if (tid == 0)
        for (int j = 0; j < N; ++j) {
            const int p = pv[j];
            if (p != j) {
                const c32 t = xs[j];
                xs[j] = xs[p];
                xs[p] = t;
            }
        }
    __syncthreads();
    constexpr int BS = 32;
    for (int kb = 0; kb < N; kb += BS) {
        if (tid < 64) {
            c32 x = lane < BS ? xs[kb + lane] : c32{0.0f, 0.0f};
            for (int j = 0; j < BS - 1; ++j) {
                const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                if (lane > j && lane < BS) {
                    const c32 l = ld(Ab, (size_t)(kb + j) * lda + kb + lane);
                    x.re -= l.re * xr - l.im * xi;
                    x.im -= l.re * xi + l.im * xr;
                }
            }
            if (lane < BS) xs[kb + lane] = x;
        }
        __syncthreads();
        for (int r = kb + BS + tid; r < N; r += 1024) {
            c32 acc = xs[r];
#pragma unroll 8
            for (int j = 0; j < BS; ++j) {
                const c32 l = ld(Ab, (size_t)(kb + j) * lda + r);
                const c32 xj = xs[kb + j];
                acc.re -= l.re * xj.re - l.im * xj.im;
                acc.im -= l.re * xj.im + l.im * xj.re;
            }
            xs[r] = acc;
        }
        __syncthreads();
    }
    for (int kb = N - BS; kb >= 0; kb -= BS) {
        if (tid < 64) {
            c32 x = lane < BS ? xs[kb + lane] : c32{0.0f, 0.0f};
            for (int j = BS - 1; j >= 0; --j) {
                if (lane == j) x = cmul(x, cinv(ld(Ab, (size_t)(kb + j) * lda + kb + j)));
                const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                if (lane < j) {
                    const c32 u = ld(Ab, (size_t)(kb + j) * lda + kb + lane);
                    x.re -= u.re * xr - u.im * xi;
                    x.im -= u.re * xi + u.im * xr;
                }
            }
            if (lane < BS) xs[kb + lane] = x;
        }
        __syncthreads();
        for (int r = tid; r < kb; r += 1024) {
            c32 acc = xs[r];
#pragma unroll 8
            for (int j = 0; j < BS; ++j) {
                const c32 u = ld(Ab, (size_t)(kb + j) * lda + r);
                const c32 xj = xs[kb + j];
                acc.re -= u.re * xj.re - u.im * xj.im;
                acc.im -= u.re * xj.im + u.im * xj.re;
            }
            xs[r] = acc;
        }
        __syncthreads();
    }
    for (int r = tid; r < N; r += 1024) st(xb, r, xs[r]);
}

// trans = 0: x <- A^{-1} x ;  trans = 2: x <- A^{-H} x.
__global__ __launch_bounds__(1024) void k_lu_solve(const float *A, int64_t lda, int64_t a_stride,
                                                   int N, const int32_t *ipiv, float *X,
                                                   int64_t ldx, int trans) {
    extern __shared__ c32 xs[];
    __shared__ c32 red[16];
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int32_t *pv = ipiv + (size_t)b * N;
    float *xb = X + (size_t)b * ldx * 2;
    for (int r = tid; r < N; r += 1024) xs[r] = ld(xb, r);
    __syncthreads();
    if (trans == 0) {
        if (tid == 0)
            for (int j = 0; j < N; ++j) {
                const int p = pv[j];
                if (p != j) {
                    const c32 t = xs[j];
                    xs[j] = xs[p];
                    xs[p] = t;
                }
            }
        __syncthreads();
        // L y = P b, unit lower, column sweep.
        for (int j = 0; j < N - 1; ++j) {
            const c32 xj = xs[j];
            for (int r = j + 1 + tid; r < N; r += 1024) {
                const c32 l = ld(Ab, (size_t)j * lda + r);
                xs[r].re -= l.re * xj.re - l.im * xj.im;
                xs[r].im -= l.re * xj.im + l.im * xj.re;
            }
            __syncthreads();
        }
        // U x = y, column sweep from the bottom.
        for (int j = N - 1; j >= 0; --j) {
            const c32 u = ld(Ab, (size_t)j * lda + j);
            const c32 xj = cmul(xs[j], cinv(u));
            __syncthreads();
            if (tid == 0) xs[j] = xj;
            for (int r = tid; r < j; r += 1024) {
                const c32 a = ld(Ab, (size_t)j * lda + r);
                xs[r].re -= a.re * xj.re - a.im * xj.im;
                xs[r].im -= a.re * xj.im + a.im * xj.re;
            }
            __syncthreads();
        }
    } else {
        // U^H y = b: y_j = (b_j - sum_{i<j} conj(U_ij) y_i) / conj(U_jj).
        for (int j = 0; j < N; ++j) {
            c32 part{0.0f, 0.0f};
            for (int i = tid; i < j; i += 1024) {
                const c32 u = cconj(ld(Ab, (size_t)j * lda + i));
                const c32 t = cmul(u, xs[i]);
                part.re += t.re;
                part.im += t.im;
            }
            const c32 s = block_sum(part, red);
            if (tid == 0) {
                const c32 d = cconj(ld(Ab, (size_t)j * lda + j));
                c32 v = xs[j];
                v.re -= s.re;
                v.im -= s.im;
                xs[j] = cmul(v, cinv(d));
            }
            __syncthreads();
        }
        // L^H z = y (unit): z_j = y_j - sum_{i>j} conj(L_ij) z_i.
        for (int j = N - 2; j >= 0; --j) {
            c32 part{0.0f, 0.0f};
            for (int i = j + 1 + tid; i < N; i += 1024) {
                const c32 l = cconj(ld(Ab, (size_t)j * lda + i));
                const c32 t = cmul(l, xs[i]);
                part.re += t.re;
                part.im += t.im;
            }
            const c32 s = block_sum(part, red);
            if (tid == 0) {
                xs[j].re -= s.re;
                xs[j].im -= s.im;
            }
            __syncthreads();
        }
        if (tid == 0)
            for (int j = N - 1; j >= 0; --j) {
                const int p = pv[j];
                if (p != j) {
                    const c32 t = xs[j];
                    xs[j] = xs[p];
                    xs[p] = t;
                }
            }
        __syncthreads();
    }
    for (int r = tid; r < N; r += 1024) st(xb, r, xs[r]);
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x over the true n x n block.
// Block = 256 rows of one system; x staged through LDS in chunks.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_residual_f64(const float *A, int64_t lda,
                                                      int64_t a_stride, int n, const float *bvec,
                                                      int64_t ldb, const double *x, int64_t ldx,
                                                      double *r, double *norms) {
    const int b = blockIdx.y;
    const int row = blockIdx.x * 256 + threadIdx.x;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const double *xb = x + (size_t)b * ldx * 2;
    __shared__ double xsh[2][256];
    double ar = 0.0, ai = 0.0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        __syncthreads();
        const int c = c0 + threadIdx.x;
        xsh[0][threadIdx.x] = c < n ? xb[2 * c] : 0.0;
        xsh[1][threadIdx.x] = c < n ? xb[2 * c + 1] : 0.0;
        __syncthreads();
        if (row < n) {
            const int cend = min(256, n - c0);
            for (int cc = 0; cc < cend; ++cc) {
                const c32 a = ld(Ab, (size_t)(c0 + cc) * lda + row);
                const double xr = xsh[0][cc], xi = xsh[1][cc];
                ar = fma((double)a.re, xr, ar);
                ar = fma(-(double)a.im, xi, ar);
                ai = fma((double)a.re, xi, ai);
                ai = fma((double)a.im, xr, ai);
            }
        }
    }
    double rn = 0.0, bn = 0.0;
    if (row < n) {
        const double br = bvec[((size_t)b * ldb + row) * 2], bi = bvec[((size_t)b * ldb + row) * 2 + 1];
        const double rr = br - ar, ri = bi - ai;
        r[((size_t)b * ldx + row) * 2] = rr;
        r[((size_t)b * ldx + row) * 2 + 1] = ri;
        rn = fmax(fabs(rr), fabs(ri));
        bn = fmax(fabs(br), fabs(bi));
    }
    for (int o = 32; o >= 1; o >>= 1) {
        rn = fmax(rn, __shfl_xor(rn, o));
        bn = fmax(bn, __shfl_xor(bn, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b),
                  static_cast<unsigned long long>(__double_as_longlong(rn)));
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b + 1),
                  static_cast<unsigned long long>(__double_as_longlong(bn)));
    }
}

__global__ void k_copy_f32_to_f64(const float *src, double *dst, int64_t ls, int64_t ld, int n) {
    const int b = blockIdx.y;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < 2 * n; i += gridDim.x * 256)
        dst[(size_t)b * ld * 2 + i] = (double)src[(size_t)b * ls * 2 + i];
}

__global__ void k_copy_f64_to_f32(const double *src, float *dst, int64_t ls, int64_t ld, int n) {
    const int b = blockIdx.y;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < 2 * n; i += gridDim.x * 256)
        dst[(size_t)b * ld * 2 + i] = (float)src[(size_t)b * ls * 2 + i];
}

__global__ void k_axpy_f32_to_f64(const float *d, double *x, int64_t ldd, int64_t ldx, int n,
                                  const int32_t *active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < 2 * n; i += gridDim.x * 256)
        x[(size_t)b * ldx * 2 + i] += (double)d[(size_t)b * ldd * 2 + i];
}

__global__ void k_rowsplit_to_col(const float *re, const float *im, int n, float *A, int64_t lda,
                                  int64_t a_stride) {
    const int c = blockIdx.x;
    const int b = blockIdx.y;
    for (int r = threadIdx.x; r < n; r += 256) {
        const size_t s = ((size_t)b * n + r) * n + c;
        st(A + (size_t)b * a_stride * 2, (size_t)c * lda + r, c32{re[s], im[s]});
    }
}

__global__ void k_col_to_rowsplit(const float *A, int64_t lda, int64_t a_stride, int n, float *re,
                                  float *im) {
    const int r = blockIdx.x;
    const int b = blockIdx.y;
    for (int c = threadIdx.x; c < n; c += 256) {
        const c32 v = ld(A + (size_t)b * a_stride * 2, (size_t)c * lda + r);
        const size_t s = ((size_t)b * n + r) * n + c;
        re[s] = v.re;
        im[s] = v.im;
    }
}

}  // namespace

void launch_pad_identity(float *A, int64_t lda, int64_t a_stride, int n, int N, int B,
                         hipStream_t stream) {
    hipLaunchKernelGGL(k_pad_identity, dim3(N, B), dim3(256), 0, stream, A, lda, a_stride, n, N);
}

void launch_norm1(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *norms,
                  hipStream_t stream) {
    launch_fill_zero_f64(norms, (size_t)B, stream);
    hipLaunchKernelGGL(k_norm1, dim3((n + 3) / 4, B), dim3(256), 0, stream, A, lda, a_stride, n,
                       norms);
}

void lu_factor(const LuWorkspace &w, hipStream_t stream) {
    static_assert(kLuPad % kNB == 0, "padded order must hold whole panels");
    // int32 zero has the bit pattern of float zero
    launch_fill_zero(reinterpret_cast<float *>(w.info), (size_t)w.B, stream);
    for (int k = 0; k < w.N; k += kNB) {
        const int nb = std::min(kNB, w.N - k);
        if (w.N <= 1024 && nb == kNB)
            hipLaunchKernelGGL(k_lu_panel_reg, dim3(w.B), dim3(1024), 0, stream, w.A, w.lda,
                               w.a_stride, w.N, k, w.ipiv, w.info);
        else
            hipLaunchKernelGGL(k_lu_panel, dim3(w.B), dim3(1024), 0, stream, w.A, w.lda,
                               w.a_stride, w.N, k, nb, w.ipiv, w.info);
        hipLaunchKernelGGL(k_lu_swap_trsm, dim3((w.N + 255) / 256, w.B), dim3(256), 0, stream, w.A,
                           w.lda, w.a_stride, w.N, k, nb, w.ipiv);
        const int rest = w.N - k - nb;
        if (rest > 0) {
            if (nb == kNB)
                hipLaunchKernelGGL(k_lu_gemm_mfma, dim3((rest + 63) / 64, (rest + 63) / 64, w.B),
                                   dim3(256), 0, stream, w.A, w.lda, w.a_stride, w.N, k);
            else
                hipLaunchKernelGGL(k_lu_gemm, dim3((rest + 63) / 64, (rest + 63) / 64, w.B),
                                   dim3(256), 0, stream, w.A, w.lda, w.a_stride, w.N, k, nb);
        }
    }
}

void lu_solve(const LuWorkspace &w, float *x, int64_t ldx, int trans, hipStream_t stream) {
    const size_t shm = sizeof(float) * 2 * (size_t)w.N;
    if (trans == 0 && w.N % 32 == 0)
        hipLaunchKernelGGL(k_lu_solve_blocked, dim3(w.B), dim3(1024), shm, stream, w.A, w.lda,
                           w.a_stride, w.N, w.ipiv, x, ldx);
    else
        hipLaunchKernelGGL(k_lu_solve, dim3(w.B), dim3(1024), shm, stream, w.A, w.lda,
                           w.a_stride, w.N, w.ipiv, x, ldx, trans);
}

void launch_residual_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                         const float *b, int64_t ldb, const double *x, int64_t ldx, double *r,
                         double *norms, hipStream_t stream) {
    launch_fill_zero_f64(norms, (size_t)2 * B, stream);
    hipLaunchKernelGGL(k_residual_f64, dim3((n + 255) / 256, B), dim3(256), 0, stream, A, lda,
                       a_stride, n, b, ldb, x, ldx, r, norms);
}

void launch_copy_f32_to_f64(const float *src, double *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream) {
    hipLaunchKernelGGL(k_copy_f32_to_f64, dim3((2 * n + 255) / 256, B), dim3(256), 0, stream, src,
                       dst, ld_src, ld_dst, n);
}

void launch_copy_f64_to_f32(const double *src, float *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream) {
    hipLaunchKernelGGL(k_copy_f64_to_f32, dim3((2 * n + 255) / 256, B), dim3(256), 0, stream, src,
                       dst, ld_src, ld_dst, n);
}

void launch_axpy_f32_to_f64(const float *d, double *x, int64_t ld_d, int64_t ld_x, int n, int B,
                            const int32_t *active, hipStream_t stream) {
    hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3((2 * n + 255) / 256, B), dim3(256), 0, stream, d, x,
                       ld_d, ld_x, n, active);
}

void launch_rowmajor_split_to_colmajor(const float *re, const float *im, int n, int B, float *A,
                                       int64_t lda, int64_t a_stride, hipStream_t stream) {
    hipLaunchKernelGGL(k_rowsplit_to_col, dim3(n, B), dim3(256), 0, stream, re, im, n, A, lda,
                       a_stride);
}

void launch_colmajor_to_rowmajor_split(const float *A, int64_t lda, int64_t a_stride, int n,
                                       int B, float *re, float *im, hipStream_t stream) {
    hipLaunchKernelGGL(k_col_to_rowsplit, dim3(n, B), dim3(256), 0, stream, A, lda, a_stride, n,
                       re, im);
}

}  // namespace hbem
