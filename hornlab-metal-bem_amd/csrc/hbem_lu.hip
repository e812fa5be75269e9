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

// Row swaps of panel [k, k+nb) on every column outside the panel, then the
// unit-lower triangular solve U12 = L11^{-1} A12 on columns >= k + nb.
// One thread per column.
__global__ __launch_bounds__(256) void k_lu_swap_trsm(float *A, int64_t lda, int64_t a_stride,
                                                      int N, int k, int nb,
                                                      const int32_t *ipiv) {
    const int b = blockIdx.y;
    float *Ab = A + (size_t)b * a_stride * 2;
    __shared__ c32 L11[kNB][kNB + 1];
    __shared__ int piv[kNB];
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e % nb, j = e / nb;  // row i, col j of the diagonal block
        L11[i][j] = ld(Ab, (size_t)(k + j) * lda + k + i);
    }
    if (threadIdx.x < nb) piv[threadIdx.x] = ipiv[(size_t)b * N + k + threadIdx.x];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= N || (c >= k && c < k + nb)) return;
    const size_t cc = (size_t)c * lda;
    for (int j = 0; j < nb; ++j) {
        const int p = piv[j];
        if (p != k + j) {
            const c32 x = ld(Ab, cc + k + j), y = ld(Ab, cc + p);
            st(Ab, cc + k + j, y);
            st(Ab, cc + p, x);
        }
    }
    if (c < k + nb) return;
    c32 x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i)
        if (i < nb) x[i] = ld(Ab, cc + k + i);
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        if (i >= nb) break;
        c32 acc = x[i];
#pragma unroll
        for (int j = 0; j < kNB; ++j) {
            if (j >= i) break;
            const c32 l = L11[i][j];
            acc.re -= l.re * x[j].re - l.im * x[j].im;
            acc.im -= l.re * x[j].im + l.im * x[j].re;
        }
        x[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < kNB; ++i)
        if (i < nb) st(Ab, cc + k + i, x[i]);
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
    // int32 zero has the bit pattern of float zero
    launch_fill_zero(reinterpret_cast<float *>(w.info), (size_t)w.B, stream);
    for (int k = 0; k < w.N; k += kNB) {
        const int nb = std::min(kNB, w.N - k);
        hipLaunchKernelGGL(k_lu_panel, dim3(w.B), dim3(1024), 0, stream, w.A, w.lda, w.a_stride,
                           w.N, k, nb, w.ipiv, w.info);
        hipLaunchKernelGGL(k_lu_swap_trsm, dim3((w.N + 255) / 256, w.B), dim3(256), 0, stream, w.A,
                           w.lda, w.a_stride, w.N, k, nb, w.ipiv);
        const int rest = w.N - k - nb;
        if (rest > 0)
            hipLaunchKernelGGL(k_lu_gemm, dim3((rest + 63) / 64, (rest + 63) / 64, w.B), dim3(256),
                               0, stream, w.A, w.lda, w.a_stride, w.N, k, nb);
    }
}

void lu_solve(const LuWorkspace &w, float *x, int64_t ldx, int trans, hipStream_t stream) {
    const size_t shm = sizeof(float) * 2 * (size_t)w.N;
    hipLaunchKernelGGL(k_lu_solve, dim3(w.B), dim3(1024), shm, stream, w.A, w.lda, w.a_stride, w.N,
                       w.ipiv, x, ldx, trans);
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
