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
// maximal |re| + |im| (icamax) and ipiv[j] is the row exchanged with row j.
// Unlike cgetrf, a step's exchanges are not applied to the factor columns
// of earlier steps: L is stored block column by block column in the row
// order of its own step, and the solves interleave the exchanges with the
// block sweeps (the same operator, fewer strided row swaps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "hbem_internal.h"

namespace hbem {
namespace {

constexpr int kNB = 32;  // panel width
static_assert(kLuGather == 4 * kNB + 4, "gather record layout");

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

// v's value in lane j of the wave (v_readlane; j uniform).
__device__ __forceinline__ float lane_value(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
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

// The same with the true order of each system in n_sys[b] (systems of
// different orders sharing one padded order N).
__global__ __launch_bounds__(256) void k_pad_identity_var(float *A, int64_t lda, int64_t a_stride,
                                                          const int32_t *n_sys, int N) {
    const int c = blockIdx.x;
    const int b = blockIdx.y;
    const int n = n_sys[b];
    float *Ab = A + (size_t)b * a_stride * 2;
    for (int r = threadIdx.x; r < N; r += 256) {
        if (r < n && c < n) continue;
        st(Ab, (size_t)c * lda + r, c32{r == c ? 1.0f : 0.0f, 0.0f});
    }
}

// norms[b] = max_c sum_r |a_rc| over the true n x n block (clange '1', FP32
// moduli summed in FP64) and amax[b] = max |a_rc|.  One wave per column;
// both must be zero on entry.
__global__ __launch_bounds__(256) void k_norm1(const float *A, int64_t lda, int64_t a_stride, int n,
                                               double *norms, double *amax) {
    // 16 columns per block (4 per wave), one atomic pair per block: the
    // per-matrix maxima are contended by every block of that matrix
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.y;
    const float *Ab = A + (size_t)b * a_stride * 2;
    double colmax = 0.0;
    float m = 0.0f;
    for (int q = 0; q < 4; ++q) {
        const int c = blockIdx.x * 16 + wave * 4 + q;
        if (c >= n) break;
        double s = 0.0;
        // two rows per lane and load (16 bytes; lda is even, so aligned),
        // loads of several iterations in flight
#pragma unroll 4
        for (int r = 2 * lane; r < n; r += 128) {
            const float4 v = *reinterpret_cast<const float4 *>(Ab + 2 * ((size_t)c * lda + r));
            const float a0 = __builtin_sqrtf(fmaf(v.x, v.x, v.y * v.y));
            const float a1 = r + 1 < n ? __builtin_sqrtf(fmaf(v.z, v.z, v.w * v.w)) : 0.0f;
            s += (double)a0 + (double)a1;
            m = fmaxf(m, fmaxf(a0, a1));
        }
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
        colmax = fmax(colmax, s);
    }
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ double red[2][4];
    if (lane == 0) {
        red[0][wave] = colmax;
        red[1][wave] = (double)m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double cm = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
        const double am = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
        atomicMax(reinterpret_cast<unsigned long long *>(norms + b),
                  static_cast<unsigned long long>(__double_as_longlong(cm)));
        atomicMax(reinterpret_cast<unsigned long long *>(amax + b),
                  static_cast<unsigned long long>(__double_as_longlong(am)));
    }
}

// ---------------------------------------------------------------------------
// Panel factorisation.  Rows are exchanged inside the panel only; columns
// left of the panel keep the row order of their own step (the solves apply
// each step's exchanges between the block sweeps), and columns right of the
// panel receive the step's exchanges in k_lu_swap_trsm through the composed
// gather list written here.
// ---------------------------------------------------------------------------

// Composes the kNB row exchanges (k + j <-> piv[j], j = 0..kNB-1, in order)
// of one panel into a gather list: position trow[m] receives the old row
// tsrc[m] for m < count <= 2 kNB.  Record layout (kGl ints): [0] count,
// [1 .. 2kNB] trow, [2kNB+1 .. 4kNB] tsrc.  Run by the 64 lanes of one wave.
__device__ void build_gather_list(const int32_t *piv, int k, int32_t *out) {
    const int lane = threadIdx.x & 63;
    int trow = k + lane, tsrc = k + lane;
    int cnt = kNB;
    for (int j = 0; j < kNB; ++j) {
        const int p = piv[j];
        if (p == k + j) continue;
        const unsigned long long hit = __ballot(lane < cnt && trow == p);
        int ip;
        if (hit != 0ull) {
            ip = __ffsll(static_cast<unsigned long long>(hit)) - 1;
        } else {
            ip = cnt++;
            if (lane == ip) {
                trow = p;
                tsrc = p;
            }
        }
        const int sj = __builtin_amdgcn_readlane(tsrc, j);  // j, ip are wave-uniform
        const int sip = __builtin_amdgcn_readlane(tsrc, ip);
        if (lane == j) tsrc = sip;
        if (lane == ip) tsrc = sj;
    }
    if (lane < cnt) {
        out[1 + lane] = trow;
        out[1 + 2 * kNB + lane] = tsrc;
    }
    if (lane == 0) out[0] = cnt;
}

// Order key of a pivot candidate: larger |re| + |im| first, then the smaller
// row (the first maximum, as icamax).  Magnitudes are >= 0, so their float
// bits order like the values.
__device__ __forceinline__ unsigned long long pivot_key(c32 v, int row) {
    const float mag = fabsf(v.re) + fabsf(v.im);
    return (static_cast<unsigned long long>(__float_as_uint(mag)) << 32) |
           static_cast<unsigned long long>(0xffffffffu - static_cast<unsigned>(row));
}

// One step of a wave maximum over 64-bit keys: each lane takes the larger of
// its key and the key of the lane the DPP control names (rows outside
// kRowMask keep their own key).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ unsigned long long key_max_dpp(unsigned long long key) {
    const int lo = static_cast<int>(static_cast<unsigned>(key));
    const int hi = static_cast<int>(static_cast<unsigned>(key >> 32));
    const unsigned olo =
        static_cast<unsigned>(__builtin_amdgcn_update_dpp(lo, lo, kCtrl, kRowMask, 0xf, false));
    const unsigned ohi =
        static_cast<unsigned>(__builtin_amdgcn_update_dpp(hi, hi, kCtrl, kRowMask, 0xf, false));
    const unsigned long long other = (static_cast<unsigned long long>(ohi) << 32) | olo;
    return other > key ? other : key;
}

// Maximum over the 16 lanes of each DPP row, left in every lane of the row:
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror.
__device__ __forceinline__ unsigned long long row_max_u64(unsigned long long key) {
    key = key_max_dpp<0xB1, 0xF>(key);
    key = key_max_dpp<0x4E, 0xF>(key);
    key = key_max_dpp<0x141, 0xF>(key);
    return key_max_dpp<0x140, 0xF>(key);
}

// Maximum over the wave, returned in every lane: the row maximum, then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3 leave
// the wave maximum in lane 63, which is read back as a uniform value.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long key) {
    key = row_max_u64(key);
    key = key_max_dpp<0x142, 0xA>(key);
    key = key_max_dpp<0x143, 0xC>(key);
    const unsigned lo = static_cast<unsigned>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(key)), 63));
    const unsigned hi = static_cast<unsigned>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(key >> 32)), 63));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// 32-bit pivot search for the register panel: a candidate's key is the bit
// pattern of its |re| + |im| with the sign bit cleared (non-negative, so
// integer order is value order; a NaN sorts above every number), -1 for a
// lane without a candidate.  One DPP step is a single integer max.
__device__ __forceinline__ int pivot_bits(c32 v) {
    return static_cast<int>(__float_as_uint(fabsf(v.re) + fabsf(v.im)) & 0x7fffffffu);
}
// (lanes outside kRowMask see INT_MIN, the identity of the max, which lets
// the compiler fold the DPP move into the max)
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int imax_dpp(int x) {
    return max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ int row_max_i32(int x) {
    x = imax_dpp<0xB1, 0xF>(x);
    x = imax_dpp<0x4E, 0xF>(x);
    x = imax_dpp<0x141, 0xF>(x);
    return imax_dpp<0x140, 0xF>(x);
}
__device__ __forceinline__ int wave_max_i32(int x) {
    x = row_max_i32(x);
    x = imax_dpp<0x142, 0xA>(x);
    x = imax_dpp<0x143, 0xC>(x);
    return __builtin_amdgcn_readlane(x, 63);
}

// Partial-pivoting elimination of the kNB panel columns over the rows held
// by the block, one row per thread: the pivot of column j is the row not yet
// chosen with the largest |re| + |im| (ties: the smallest id), and every
// other unchosen row is eliminated against it.  Returns the column the
// thread's row became the pivot of (or -1); v ends as that row's factored
// values (multipliers left of its pivot column, U entries from it).  win_s[j]
// is the id of column j's pivot (-1 without candidates), mag_s[j] its
// |re| + |im|.
__device__ int eliminate_rows(c32 (&v)[kNB], bool own, unsigned id,
                              unsigned long long *best_s, c32 *prow, int *win_s,
                              float *mag_s) {
    const int tid = threadIdx.x, lane = tid & 63;
    int rank = -1;
    if (tid == 0) {
        best_s[0] = 0ull;
        best_s[1] = 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kNB; ++j) {
        unsigned long long key = (own && rank < 0) ? pivot_key(v[j], (int)id) : 0ull;
        key = wave_max_u64(key);
        if (lane == 0) atomicMax(&best_s[j & 1], key);
        __syncthreads();
        const unsigned long long best = best_s[j & 1];
        const unsigned wid = 0xffffffffu - static_cast<unsigned>(best);
        const bool holder = best != 0ull && own && rank < 0 && id == wid;
        if (holder) {
            rank = j;
#pragma unroll
            for (int c = j; c < kNB; ++c) prow[c] = v[c];
        }
        if (tid == 0) {
            best_s[(j + 1) & 1] = 0ull;
            win_s[j] = best != 0ull ? static_cast<int>(wid) : -1;
            mag_s[j] = __uint_as_float(static_cast<unsigned>(best >> 32));
        }
        __syncthreads();
        if (best != 0ull && own && rank < 0) {
            const c32 piv = prow[j];
            const bool nz = (piv.re != 0.0f || piv.im != 0.0f);
            c32 l = v[j];
            if (nz) l = cmul(l, cinv(piv));
            v[j] = l;
#pragma unroll
            for (int c = j + 1; c < kNB; ++c) {
                const c32 u = prow[c];
                v[c].re = fmaf(-l.re, u.re, fmaf(l.im, u.im, v[c].re));
                v[c].im = fmaf(-l.re, u.im, fmaf(-l.im, u.re, v[c].im));
            }
        }
    }
    return rank;
}

// Panels with more than 1024 rows: tournament pivoting (communication-
// avoiding LU).  Step 1: every 1024-row chunk runs the elimination on its
// rows and nominates its kNB pivot rows.  cand[b][g * kNB + j].
__global__ __launch_bounds__(1024) void k_lu_panel_nominate(const float *A, int64_t lda,
                                                            int64_t a_stride, int N, int k,
                                                            int32_t *cand) {
    const int g = blockIdx.x, b = blockIdx.y, G = gridDim.x;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int r = k + g * 1024 + (int)threadIdx.x;
    const bool own = r < N;
    c32 v[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) v[c] = own ? ld(Ab, (size_t)(k + c) * lda + r) : c32{0.f, 0.f};
    __shared__ unsigned long long best_s[2];
    __shared__ c32 prow[kNB];
    __shared__ int win_s[kNB];
    __shared__ float mag_s[kNB];
    (void)eliminate_rows(v, own, (unsigned)r, best_s, prow, win_s, mag_s);
    if (threadIdx.x < kNB) cand[(size_t)b * G * kNB + g * kNB + threadIdx.x] = win_s[threadIdx.x];
}

// Step 2, one block per system: the elimination over the G * kNB nominated
// rows (original values) picks the panel's kNB pivot rows and factors the
// top block.  Writes ipiv (the row exchanges in order), the gather list, the
// factored top block top[b][kNB][kNB] (row j = pivot j) and the original
// panel values of the gathered source rows gsrc[b][2 kNB][kNB].
__global__ __launch_bounds__(1024) void k_lu_panel_tournament(
    const float *A, int64_t lda, int64_t a_stride, int N, int k, int G, const int32_t *cand,
    int32_t *ipiv, int32_t *info, int32_t *gl, int64_t gl_stride, float *top, float *gsrc) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int ncand = G * kNB;
    const int id = tid < ncand ? cand[(size_t)b * ncand + tid] : -1;
    const bool own = id >= 0;
    c32 v[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) v[c] = own ? ld(Ab, (size_t)(k + c) * lda + id) : c32{0.f, 0.f};
    __shared__ unsigned long long best_s[2];
    __shared__ c32 prow[kNB];
    __shared__ int win_s[kNB];
    __shared__ float mag_s[kNB];
    __shared__ int piv_s[kNB];
    const int rank = eliminate_rows(v, own, (unsigned)id, best_s, prow, win_s, mag_s);
    float *topb = top + (size_t)b * kNB * kNB * 2;
    if (rank >= 0) {
#pragma unroll
        for (int c = 0; c < kNB; ++c) st(topb, (size_t)rank * kNB + c, v[c]);
    }
    if (wave == 0) {
        // exchanges in order: the winner of column j moves to row k + j.
        // One touched position per lane: lane i < nt holds position tpos,
        // whose row is tcon; lanes 0..31 are positions k..k+31, and a
        // position outside them joins the first free lane when a winner
        // is taken from it (untouched positions hold their own row).
        int tpos = lane < kNB ? k + lane : -1, tcon = tpos;
        int nt = kNB, bad = 0;
        for (int j = 0; j < kNB; ++j) {
            const int w = win_s[j] >= 0 ? win_s[j] : k + j;
            const unsigned long long at = __ballot(lane < nt && tcon == w);
            int lp;  // lane of the current position of row w
            if (at != 0ull) {
                lp = __ffsll(static_cast<unsigned long long>(at)) - 1;
            } else {
                lp = nt++;
                if (lane == lp) tpos = tcon = w;
            }
            const int p = __builtin_amdgcn_readlane(tpos, lp);
            const int cj = __builtin_amdgcn_readlane(tcon, j);  // position k + j is lane j
            const int cp = __builtin_amdgcn_readlane(tcon, lp);
            if (lane == j) tcon = cp;
            if (lane == lp) tcon = cj;
            if (lane == 0) {
                piv_s[j] = p;
                ipiv[(size_t)b * N + k + j] = p;
            }
            if (!(mag_s[j] > 0.0f) && bad == 0) bad = k + j + 1;
        }
        if (lane == 0 && bad != 0 && info[b] == 0) info[b] = bad;
    }
    __syncthreads();
    int32_t *g = gl + (size_t)b * gl_stride;
    if (wave == 0) build_gather_list(piv_s, k, g);
    __syncthreads();
    const int cnt = g[0];
    float *gs = gsrc + (size_t)b * 2 * kNB * kNB * 2;
    for (int e = tid; e < cnt * kNB; e += (int)blockDim.x) {
        const int m = e / kNB, c = e % kNB;
        st(gs, (size_t)m * kNB + c, ld(Ab, (size_t)(k + c) * lda + g[1 + 2 * kNB + m]));
    }
}

// Step 3: the panel rows in their new order: rows k..k+kNB-1 get the
// factored top block, every row below gets its multipliers
// l = a U11^{-1} (a row-vector triangular solve), where a is the row's
// panel values after the exchanges.  One thread per row.
__global__ __launch_bounds__(256) void k_lu_panel_apply(float *A, int64_t lda, int64_t a_stride,
                                                        int N, int k, const int32_t *gl,
                                                        int64_t gl_stride, const float *top,
                                                        const float *gsrc) {
    const int b = blockIdx.y;
    float *Ab = A + (size_t)b * a_stride * 2;
    const float *topb = top + (size_t)b * kNB * kNB * 2;
    const float *gs = gsrc + (size_t)b * 2 * kNB * kNB * 2;
    const int32_t *g = gl + (size_t)b * gl_stride;
    __shared__ c32 U[kNB][kNB + 1];
    __shared__ c32 rdiag[kNB];
    __shared__ int trow[2 * kNB];
    __shared__ int cnt_s;
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) U[e / kNB][e % kNB] = ld(topb, e);
    if (tid == 0) cnt_s = g[0];
    __syncthreads();
    const int cnt = cnt_s;
    if (tid < cnt) trow[tid] = g[1 + tid];
    if (tid < kNB) {
        const c32 d = U[tid][tid];
        rdiag[tid] = (d.re != 0.0f || d.im != 0.0f) ? cinv(d) : c32{0.0f, 0.0f};
    }
    __syncthreads();
    const int r = k + blockIdx.x * 256 + tid;
    if (r >= N) return;
    if (r < k + kNB) {
#pragma unroll
        for (int c = 0; c < kNB; ++c) st(Ab, (size_t)(k + c) * lda + r, U[r - k][c]);
        return;
    }
    int m = -1;
    for (int i = 0; i < cnt; ++i)
        if (trow[i] == r) m = i;
    c32 x[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        x[c] = m >= 0 ? ld(gs, (size_t)m * kNB + c) : ld(Ab, (size_t)(k + c) * lda + r);
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
        c32 acc = x[c];
#pragma unroll
        for (int i = 0; i < c; ++i) {
            const c32 u = U[i][c];
            acc.re = fmaf(-x[i].re, u.re, fmaf(x[i].im, u.im, acc.re));
            acc.im = fmaf(-x[i].re, u.im, fmaf(-x[i].im, u.re, acc.im));
        }
        x[c] = cmul(acc, rdiag[c]);
    }
#pragma unroll
    for (int c = 0; c < kNB; ++c) st(Ab, (size_t)(k + c) * lda + r, x[c]);
}

// Register panel for N - k <= 1024, one wave per 64 rows (blockDim.x =
// 64 ceil((N - k) / 64)): thread t holds the kNB panel values of one row,
// the column loop is unrolled so that every register index is
// static (a rolled loop selecting the column by predicates issues about
// twice the instructions per column and runs twice as long; the kernel is
// bound by the dependent instruction stream of a wave per column, not by
// its barriers or its wave count).  A row exchange swaps only the logical row numbers of the two
// threads involved (no data moves); every thread writes its values back to
// its final logical row.  One block barrier per column: each wave reduces
// its pivot candidate with DPP over 32-bit magnitude keys (one integer max
// per step; the smallest row among equal magnitudes takes a second
// reduction only when the ballot shows a tie), and its candidate lane
// publishes the key, its row and the row's values from the column on to a
// double-buffered per-wave slot; after the barrier every wave reduces the
// 16 keys (one DPP row) and reads the winning slot.  A slot of
// column j is rewritten at column j + 2, after every reader has passed the
// barrier of column j + 1.
__global__ __launch_bounds__(1024) void k_lu_panel_reg(float *A, int64_t lda, int64_t a_stride,
                                                       int N, int k, int32_t *ipiv,
                                                       int32_t *info, int32_t *gl,
                                                       int64_t gl_stride) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    float *Ab = A + (size_t)b * a_stride * 2;
    const bool own = k + tid < N;
    int myrow = k + tid;  // logical row held by this thread
    c32 v[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        v[c] = own ? ld(Ab, (size_t)(k + c) * lda + myrow) : c32{0.0f, 0.0f};
    __shared__ int mag_s[2][16];  // a wave's largest candidate key (-1: none)
    __shared__ int rid_s[2][16];  // the logical row of that wave's candidate
    __shared__ c32 row_s[2][16][kNB];
    __shared__ int piv_s[kNB];
    // the block has one wave per 64 rows; slots of absent waves hold no key
    if (tid < 32 && (tid & 15) >= (int)(blockDim.x >> 6)) mag_s[tid >> 4][tid & 15] = -1;
    int bad = 0;  // thread 0: first column without a nonzero pivot, plus one
#pragma unroll
    for (int j = 0; j < kNB; ++j) {
        const int col = k + j, slot = j & 1;
        const bool cand = own && myrow >= col;
        const int key = cand ? pivot_bits(v[j]) : -1;
        const int wkey = wave_max_i32(key);
        // the wave's candidate: its lane of largest key, of those the
        // smallest row (a second reduction only for a tie)
        const bool top = cand && key == wkey;
        const unsigned long long hit = __ballot(top);
        bool win = top;
        if (__popcll(hit) > 1) {
            const int t = wave_max_i32(top ? 0x7fffffff - myrow : -1);
            win = top && 0x7fffffff - myrow == t;
        }
        if (win) {
            rid_s[slot][wave] = myrow;
#pragma unroll
            for (int c = j; c < kNB; ++c) row_s[slot][wave][c] = v[c];
        }
        if (lane == 0) mag_s[slot][wave] = wkey;
        __syncthreads();
        const int mine = mag_s[slot][lane & 15];
        const int best = row_max_i32(mine);
        const unsigned hw = static_cast<unsigned>(__ballot(lane < 16 && mine == best));
        int wbest = __builtin_ctz(hw);
        if (__popc(hw) > 1) {
            // tied waves: the smallest row among their candidates
            const int rk = (lane < 16 && mine == best) ? 0x7fffffff - rid_s[slot][lane] : -1;
            const int t = row_max_i32(rk);
            wbest = __builtin_ctz(static_cast<unsigned>(__ballot(lane < 16 && rk == t)));
        }
        const int p = rid_s[slot][wbest];
        if (tid == 0) {
            piv_s[j] = p;
            if (!(__int_as_float(best) > 0.0f) && bad == 0) bad = col + 1;
        }
        // exchange logical rows col <-> p
        if (own && myrow == p)
            myrow = col;
        else if (own && myrow == col)
            myrow = p;
        // rows below the pivot: l = a / pivot, a_c -= l u_c
        if (own && myrow > col) {
            const c32 *prow = row_s[slot][wbest];
            const c32 piv = prow[j];
            c32 l = v[j];
            if (piv.re != 0.0f || piv.im != 0.0f) {
                const float s = 1.0f / (piv.re * piv.re + piv.im * piv.im);
                l = cmul(l, c32{piv.re * s, -piv.im * s});
            }
            v[j] = l;
#pragma unroll
            for (int c = j + 1; c < kNB; ++c) {
                const c32 u = prow[c];
                v[c].re = fmaf(-l.re, u.re, fmaf(l.im, u.im, v[c].re));
                v[c].im = fmaf(-l.re, u.im, fmaf(-l.im, u.re, v[c].im));
            }
        }
    }
    if (own) {
#pragma unroll
        for (int c = 0; c < kNB; ++c) st(Ab, (size_t)(k + c) * lda + myrow, v[c]);
    }
    if (tid == 0 && bad != 0 && info[b] == 0) info[b] = bad;
    __syncthreads();
    // the step's pivots, once after the loop
    if (tid < kNB) ipiv[(size_t)b * N + k + tid] = piv_s[tid];
    if (wave == 0) build_gather_list(piv_s, k, gl + (size_t)b * gl_stride);
}

// Block row of U for the trailing columns of an outer block [K0, K1 =
// K0 + kLuOuter): the exchanges of the block's kLuOuter / kNB steps applied
// to columns [K1, N) (rows [K0, N)), then U12 = L11^{-1} A12 with the
// block's unit-lower L11, kBrCols columns per workgroup, the block row held
// in LDS throughout.  One launch in place of the exchange launch and the
// per-step solve and rank-kNB update launches.  The steps' gather records
// are composed into one source row per row (map_s); the rows below K1 that
// the exchanges move are read before, and written after, one barrier.  The
// workgroups of one system share an XCD (B % 8 == 0), so L11 is read into
// one L2.
constexpr int kBrCols = 32;
constexpr int kBrMaxOrder = 8192;  // map_s: N - K0 ints of dynamic LDS
// A workgroup's time is set by its serial steps (about 160 us at 256 x 32),
// so the launch pays when few rounds of workgroups fill the chip; beyond
// this many workgroups the per-step launches run faster (measured: 40
// systems of order 1088 gain, 64 of order 8000 lose).
constexpr int kBrMaxGroups = 1536;
__global__ __launch_bounds__(256, 2) void k_lu_block_row(float *A, int64_t lda, int64_t a_stride,
                                                      int N, int K0, int K1, const int32_t *gl,
                                                      int64_t gl_stride) {
    extern __shared__ int map_s[];          // [N - K0] source row of row K0 + i
    __shared__ c32 X[kLuOuter][kBrCols];    // the block row, column c of row r at c ^ (r & 31)
    __shared__ int moved_s[kLuOuter];
    __shared__ int nmoved_s;
    const int per = gridDim.x, B = gridDim.y;
    int b = blockIdx.y, tile = blockIdx.x;
    if ((B & 7) == 0) {
        const int L = blockIdx.x + blockIdx.y * per, k = L >> 3;
        b = (L & 7) + 8 * (k / per);
        tile = k % per;
    }
    const int tid = threadIdx.x;
    const int c0 = K1 + tile * kBrCols;
    const int nc = min(kBrCols, N - c0);
    float *Ab = A + (size_t)b * a_stride * 2;
    const int M = N - K0;
    for (int i = tid; i < M; i += 256) map_s[i] = K0 + i;
    if (tid == 0) nmoved_s = 0;
    __syncthreads();
    const int32_t *g0 = gl + (size_t)b * gl_stride + (size_t)(K0 / kNB) * kLuGather;
    for (int s = 0; s < kLuOuter / kNB; ++s) {
        const int32_t *g = g0 + (size_t)s * kLuGather;
        const int cnt = g[0];
        int dst = 0, src = 0;
        if (tid < cnt) {
            dst = g[1 + tid];
            src = map_s[g[1 + 2 * kNB + tid] - K0];
        }
        __syncthreads();
        if (tid < cnt) map_s[dst - K0] = src;
        __syncthreads();
    }
    // rows below the block that receive another row (at most kNB per step)
    for (int i = kLuOuter + tid; i < M; i += 256)
        if (map_s[i] != K0 + i) {
            const int q = atomicAdd(&nmoved_s, 1);
            if (q < kLuOuter) moved_s[q] = i;
        }
    // thread = row of the block row (256 threads, kLuOuter rows), all loads
    // of the block row and of the moved rows in flight together
    static_assert(kLuOuter == 256, "one block-row row per thread");
    __syncthreads();
    const int nm = min(nmoved_s, kLuOuter);
    {
        const int src = map_s[tid];
        const int mrow = tid < nm ? moved_s[tid] : 0;
        const int msrc = tid < nm ? map_s[mrow] : 0;
        const float *col0 = Ab + 2 * (size_t)c0 * lda;
        c32 v[kBrCols], mv[kBrCols];
#pragma unroll
        for (int c = 0; c < kBrCols; ++c)
            v[c] = c < nc ? ld(col0, (size_t)c * lda + src) : c32{0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < kBrCols; ++c) X[tid][c ^ (tid & 31)] = v[c];
#pragma unroll
        for (int c = 0; c < kBrCols; ++c)
            mv[c] = (tid < nm && c < nc) ? ld(col0, (size_t)c * lda + msrc) : c32{0.0f, 0.0f};
        __syncthreads();  // every old value this workgroup needs has been read
        if (tid < nm) {
#pragma unroll
            for (int c = 0; c < kBrCols; ++c)
                if (c < nc) st(Ab, (size_t)(c0 + c) * lda + K0 + mrow, mv[c]);
        }
    }
    for (int r0 = 0; r0 < kLuOuter; r0 += kNB) {
        // the step's rows: unit-lower solve in wave 0, lane = column; lane i
        // also holds row i of the diagonal block of L (loads coalesced along
        // the column, all in flight), whose entries reach every lane by
        // v_readlane
        if (tid < 64) {
            c32 lrow[kNB], x[kNB];
#pragma unroll
            for (int j = 0; j < kNB; ++j)
                lrow[j] = (j < tid && tid < kNB)
                              ? ld(Ab, (size_t)(K0 + r0 + j) * lda + K0 + r0 + tid)
                              : c32{0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < kNB; ++i)
                x[i] = tid < kBrCols ? X[r0 + i][tid ^ i] : c32{0.0f, 0.0f};
#pragma unroll
            for (int j = 0; j < kNB; ++j) {
                // an empty asm that ties column j of L to the final x_j: the
                // step's lane reads cannot be hoisted above it (hoisted
                // together they exhaust the scalar registers)
                asm volatile("" : "+v"(lrow[j].re), "+v"(lrow[j].im) : "v"(x[j].re), "v"(x[j].im));
#pragma unroll
                for (int i = j + 1; i < kNB; ++i) {
                    const float lre = lane_value(lrow[j].re, i), lim = lane_value(lrow[j].im, i);
                    x[i].re = fmaf(-lre, x[j].re, fmaf(lim, x[j].im, x[i].re));
                    x[i].im = fmaf(-lre, x[j].im, fmaf(-lim, x[j].re, x[i].im));
                }
            }
            if (tid < kBrCols) {
#pragma unroll
                for (int i = 0; i < kNB; ++i) X[r0 + i][tid ^ i] = x[i];
            }
        }
        __syncthreads();
        // rows below the step within the block: X_r -= L_r,step X_step;
        // thread = 4 columns x one row of each 32-row group
        const int rb = r0 + kNB, ngrp = (kLuOuter - rb) / 32;
        if (ngrp > 0) {
            constexpr int kG = kLuOuter / 32 - 1;
            const int rr = tid >> 3, cg = tid & 7;
            c32 acc[kG][4];
#pragma unroll
            for (int m = 0; m < kG; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[m][q] = m < ngrp ? X[rb + rr + 32 * m][(4 * cg + q) ^ rr] : c32{0.0f, 0.0f};
            // L of the thread's rows, kJb columns of loads in flight at once
            constexpr int kJb = 4;
            const size_t lc = (size_t)(K0 + r0) * lda + K0 + rb + rr;
            for (int j0 = 0; j0 < kNB; j0 += kJb) {
                c32 lv[kJb][kG];
#pragma unroll
                for (int jj = 0; jj < kJb; ++jj)
#pragma unroll
                    for (int m = 0; m < kG; ++m)
                        lv[jj][m] = m < ngrp ? ld(Ab, lc + (size_t)(j0 + jj) * lda + 32 * m)
                                             : c32{0.0f, 0.0f};
#pragma unroll
                for (int jj = 0; jj < kJb; ++jj) {
                    const int j = j0 + jj;
                    c32 u[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) u[q] = X[r0 + j][(4 * cg + q) ^ j];
#pragma unroll
                    for (int m = 0; m < kG; ++m) {
                        const c32 l = lv[jj][m];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            acc[m][q].re =
                                fmaf(-l.re, u[q].re, fmaf(l.im, u[q].im, acc[m][q].re));
                            acc[m][q].im =
                                fmaf(-l.re, u[q].im, fmaf(-l.im, u[q].re, acc[m][q].im));
                        }
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < kG; ++m)
                if (m < ngrp)
#pragma unroll
                    for (int q = 0; q < 4; ++q) X[rb + rr + 32 * m][(4 * cg + q) ^ rr] = acc[m][q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < kBrCols; ++c)
        if (c < nc) st(Ab, (size_t)(c0 + c) * lda + K0 + tid, X[tid][c ^ (tid & 31)]);
}

// Row exchanges and/or the unit-lower solve U12 = L11^{-1} A12 of panel
// steps on two column ranges, 64 columns per block: blocks
// [0, nblk(a)) take range [a0, a1) with mode_a, the rest [b0, b1) with
// mode_b.  mode 0: the exchanges of the n_steps consecutive steps from
// column k (their gather lists, applied in order); 1: the exchanges of step
// k, then the solve with its diagonal block; 2: the solve only.  Rows are
// staged in LDS (loads coalesced along each column), lane = column for the
// solve.
__global__ __launch_bounds__(256) void k_lu_swap_trsm(float *A, int64_t lda, int64_t a_stride,
                                                      int k, int n_steps, int a0, int a1,
                                                      int mode_a, int b0, int b1, int mode_b,
                                                      const int32_t *gl, int64_t gl_stride) {
    const int b = blockIdx.y;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int nblk_a = (a1 - a0 + 63) / 64;
    const bool in_a = (int)blockIdx.x < nblk_a;
    const int mode = in_a ? mode_a : mode_b;
    const int cb = in_a ? a0 + blockIdx.x * 64 : b0 + (blockIdx.x - nblk_a) * 64;
    const int ncols = min(64, (in_a ? a1 : b1) - cb);
    __shared__ c32 L11[kNB][kNB + 1];
    __shared__ c32 tile[2 * kNB][64 + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int steps = (mode == 0) ? n_steps : 1;
    if (mode != 0)
        for (int e = tid; e < kNB * kNB; e += 256) {
            const int i = e % kNB, j = e / kNB;
            L11[i][j] = ld(Ab, (size_t)(k + j) * lda + k + i);
        }
    for (int st_i = 0; st_i < steps; ++st_i) {
        const int ks = k + kNB * st_i;
        const int32_t *g = gl + (size_t)b * gl_stride + (size_t)st_i * kLuGather;
        // the step's gather record, lane m = entry m (wave-uniform count)
        const int cnt = (mode == 2) ? kNB : g[0];
        const bool has = lane < cnt;
        const int my_row = (mode == 2) ? ks + lane : (has ? g[1 + lane] : 0);
        const int my_src = (mode == 2) ? ks + lane : (has ? g[1 + 2 * kNB + lane] : 0);
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int cc = wave + 4 * it;
            if (cc < ncols && has) tile[lane][cc] = ld(Ab, (size_t)(cb + cc) * lda + my_src);
        }
        __syncthreads();
        if (mode != 0 && tid < ncols) {
            c32 x[kNB];
#pragma unroll
            for (int i = 0; i < kNB; ++i) x[i] = tile[i][tid];
#pragma unroll
            for (int i = 1; i < kNB; ++i) {
                c32 acc = x[i];
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    const c32 l = L11[i][j];
                    acc.re = fmaf(-l.re, x[j].re, fmaf(l.im, x[j].im, acc.re));
                    acc.im = fmaf(-l.re, x[j].im, fmaf(-l.im, x[j].re, acc.im));
                }
                x[i] = acc;
            }
#pragma unroll
            for (int i = 0; i < kNB; ++i) tile[i][tid] = x[i];
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int cc = wave + 4 * it;
            if (cc < ncols && has) st(Ab, (size_t)(cb + cc) * lda + my_row, tile[lane][cc]);
        }
        __syncthreads();
    }
}

// Rank-kNB update C -= L U on MFMA (FP32 operands, v_mfma_f32_16x16x4_f32,
// FP32 accumulation) over rows [r_begin, r_end) and columns [c_begin,
// c_end), with L = A[rows][k0, k0 + kNB) and U = A[k0, k0 + kNB)[cols]
// (both outside the C block).  64 x 64 complex output tile per block; each
// of the 4 waves owns a 32 x 32 sub-tile as 2 x 2 MFMA tiles with separate
// real and imaginary accumulators (re += Lr Ur - Li Ui, im += Lr Ui + Li Ur).
// The MFMA computes the transposed tile C^T = U^T L^T so that the 16 lanes
// of an output row group cover 16 consecutive matrix rows (coalesced
// column-major stores).  L and U are staged in LDS planes [k][row | col]
// with an 80-float stride (conflict-free reads of the 4 k-rows a wave
// touches per instruction); the thread's 16 C entries are loaded first, so
// their latency overlaps the staging and the MFMAs.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kGS = 80;  // FP32 plane stride (floats)

__device__ __forceinline__ unsigned short f32_to_bf16_rne(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16);  // inf/nan
    const unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(r >> 16);
}

struct GemmLds {
    float Lr[kNB][kGS], Li[kNB][kGS], Ur[kNB][kGS], Ui[kNB][kGS];
};

__global__ __launch_bounds__(256) void k_lu_gemm(float *A, int64_t lda, int64_t a_stride,
                                                 int r_begin, int r_end, int c_begin, int c_end,
                                                 int k0) {
    const int b = blockIdx.z;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int r0 = r_begin + blockIdx.x * 64;
    const int c0 = c_begin + blockIdx.y * 64;
    __shared__ GemmLds sm;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    c32 cv[2][2][4];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = c0 + wc + 16 * ti + 4 * lk + v;
                const int row = r0 + wr + 16 * tj + li;
                cv[ti][tj][v] = (row < r_end && col < c_end) ? ld(Ab, (size_t)col * lda + row)
                                                             : c32{0.f, 0.f};
            }
    f32x4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            accr[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            acci[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
    for (int it = 0; it < (kNB * 64) / 256; ++it) {
        const int e = tid + 256 * it;
        const int kk = e / 64, rr = e % 64;
        const c32 v = (r0 + rr < r_end) ? ld(Ab, (size_t)(k0 + kk) * lda + r0 + rr) : c32{0.f, 0.f};
        sm.Lr[kk][rr] = v.re;
        sm.Li[kk][rr] = v.im;
    }
#pragma unroll
    for (int it = 0; it < (kNB * 64) / 256; ++it) {
        const int e = tid + 256 * it;
        const int cc = e / kNB, kk = e % kNB;
        const c32 v = (c0 + cc < c_end) ? ld(Ab, (size_t)(c0 + cc) * lda + k0 + kk) : c32{0.f, 0.f};
        sm.Ur[kk][cc] = v.re;
        sm.Ui[kk][cc] = v.im;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kNB; ks += 4) {
        const int kk = ks + lk;
        float ur[2], ui[2], lr[2], lim[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            ur[t] = sm.Ur[kk][wc + 16 * t + li];
            ui[t] = sm.Ui[kk][wc + 16 * t + li];
            lr[t] = sm.Lr[kk][wr + 16 * t + li];
            lim[t] = sm.Li[kk][wr + 16 * t + li];
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
                if (row < r_end && col < c_end) {
                    c32 x = cv[ti][tj][v];
                    x.re -= accr[ti][tj][v];
                    x.im -= acci[ti][tj][v];
                    st(Ab, (size_t)col * lda + row, x);
                }
            }
}

// Rank-kLuOuter trailing update in FP16 / BF16 on the gfx950
// v_mfma_f32_16x16x32_{f16,bf16}: 128 x 128 complex outputs per block of
// 4 waves (2 x 2, 64 x 64 each = 4 x 4 MFMA tiles with separate real and
// imaginary FP32 accumulators), k_lu_gemm_trp below.  As in k_lu_gemm the
// MFMA computes the transposed tile C^T = U^T L^T, so 16 lanes of an output
// row group hold 16 consecutive rows of one column (coalesced column-major
// stores).  The operands are rounded once per outer block into packed
// 16-bit planes (k_lu_pack_l / k_lu_pack_u); the inner range is staged 32 k
// at a time: each thread loads its share of the next stage into registers
// while the MFMAs of the current stage run, then stores it into the LDS
// planes [row | col][k] (row stride 40 elements, so the 16-byte operand
// reads of a wave are conflict-free).  kPrec 1: U is scaled by the power of
// two scale[b] before rounding to FP16 and the product by 1/scale[b];
// kPrec 2: BF16, round to nearest even.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kTrTile = 128;  // output rows and columns per block
constexpr int kTrK = 32;      // k per stage
constexpr int kTrS = 40;      // LDS plane row stride (16-bit elements)

// Tile of a trailing-update block, remapped for the L2 of each XCD.  The
// hardware hands workgroups to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2: linear id g goes to XCD g % 8, so the remap gives XCD j
// the j-th contiguous eighth of the tile list (a bijection onto all tiles),
// and that list walks each system in row groups of 8 tiles, column by
// column, so that the group's 8 L strips and the current U strip stay in the
// XCD's L2 while the blocks in flight sweep the columns.  Returns the tile's
// row and column index and its system.
struct TrTile {
    int x, y, b;
};
__device__ __forceinline__ TrTile trailing_tile() {
    const int X = gridDim.x, Y = gridDim.y;
    const int total = X * Y * (int)gridDim.z;
    const int g = blockIdx.x + X * (blockIdx.y + Y * blockIdx.z);
    const int per = total / 8, rem = total % 8;
    const int xcd = g % 8, k = g / 8;
    const int t = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k;
    const int b = t / (X * Y);
    const int u = t - b * (X * Y);
    const int rg = u / (8 * Y);
    const int hx = min(8, X - 8 * rg);
    const int v = u - rg * 8 * Y;
    return TrTile{8 * rg + v % hx, v / hx, b};
}

struct __attribute__((aligned(16))) TrLds {
    unsigned short Lr[kTrTile][kTrS], Li[kTrTile][kTrS], Ur[kTrTile][kTrS], Ui[kTrTile][kTrS];
};

// C -= scale * (accr + i acci) on a wave's 64 x 64 complex block of a
// 128-tile (column base cb, row base rb; entries at or beyond c_end / r_end
// untouched).  MFMA accumulator [ti][tj][v] holds column cb + 16 ti +
// 4 (lane / 16) + v, row rb + 16 tj + lane % 16.  The loads of 32 entries are
// issued before any of their stores: the compiler cannot reorder a load past
// a store it cannot prove disjoint, so a load-modify-store per entry would
// wait one memory latency per entry.  The C loads and stores are
// non-temporal: each trailing entry is touched once per update, and keeping
// it out of L2 leaves L2 to the L and U strips that neighbouring tiles share.
__device__ __forceinline__ void tile_epilogue(float *Ab, int64_t lda, int cb, int rb, int c_end,
                                              int r_end, int lane, const f32x4 (&accr)[4][4],
                                              const f32x4 (&acci)[4][4], float scale) {
    const int li = lane & 15, lc = (lane >> 4) * 4;
#pragma unroll
    for (int t2 = 0; t2 < 4; t2 += 2) {
        c32 cv[2][4][4];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int col = cb + 16 * (t2 + h) + lc + v, row = rb + 16 * tj + li;
                    typedef float f2v __attribute__((ext_vector_type(2)));
                    f2v t = {0.0f, 0.0f};
                    if (row < r_end && col < c_end)
                        t = __builtin_nontemporal_load(
                            reinterpret_cast<const f2v *>(Ab + 2 * ((size_t)col * lda + row)));
                    cv[h][tj][v] = c32{t.x, t.y};
                }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int col = cb + 16 * (t2 + h) + lc + v, row = rb + 16 * tj + li;
                    if (row < r_end && col < c_end) {
                        c32 x = cv[h][tj][v];
                        x.re -= accr[t2 + h][tj][v] * scale;
                        x.im -= acci[t2 + h][tj][v] * scale;
                        typedef float f2v __attribute__((ext_vector_type(2)));
                        const f2v o = {x.re, x.im};
                        __builtin_nontemporal_store(
                            o, reinterpret_cast<f2v *>(Ab + 2 * ((size_t)col * lda + row)));
                    }
                }
    }
}

template <int kPrec>
__device__ __forceinline__ unsigned short tr_round(float x) {
    if constexpr (kPrec == 1) {
        const _Float16 h = static_cast<_Float16>(x);
        return *reinterpret_cast<const unsigned short *>(&h);
    } else {
        return f32_to_bf16_rne(x);
    }
}

// Packed 16-bit operands of a rank-kLuOuter trailing update, written once
// per outer block [K0, K0 + kLuOuter) so that the trailing tiles read 4 bytes
// per complex operand instead of 8 and convert nothing.  Per system, four
// planes of [N][kLuOuter] 16-bit values (re/im of L, re/im of U), k
// contiguous: L plane row i = A[K1 + i][K0 + k]; U plane row j =
// scale[b] * A[K0 + k][K1 + j] (FP16) or A[K0 + k][K1 + j] (BF16).
template <int kPrec>
__global__ __launch_bounds__(256) void k_lu_pack_l(const float *A, int64_t lda, int64_t a_stride,
                                                   int K0, int K1, int M, uint16_t *pack,
                                                   int64_t pstride, int64_t plane) {
    // 64 rows x 32 k per block: thread = (row, 8 consecutive k)
    const int b = blockIdx.z;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int row = blockIdx.x * 64 + (threadIdx.x & 63);
    const int k = blockIdx.y * 32 + (threadIdx.x >> 6) * 8;
    if (row >= M) return;
    typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
    u16x8 re, im;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const c32 v = ld(Ab, (size_t)(K0 + k + e) * lda + K1 + row);
        re[e] = tr_round<kPrec>(v.re);
        im[e] = tr_round<kPrec>(v.im);
    }
    uint16_t *pb = pack + (size_t)b * pstride + (size_t)row * kLuOuter + k;
    *reinterpret_cast<u16x8 *>(pb) = re;
    *reinterpret_cast<u16x8 *>(pb + plane) = im;
}

template <int kPrec>
__global__ __launch_bounds__(256) void k_lu_pack_u(const float *A, int64_t lda, int64_t a_stride,
                                                   int K0, int K1, int M, uint16_t *pack,
                                                   int64_t pstride, int64_t plane,
                                                   const float *scale) {
    // 8 columns x 256 k per block: thread = (column, 8 consecutive k)
    const int b = blockIdx.y;
    const float *Ab = A + (size_t)b * a_stride * 2;
    const int col = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int k = (threadIdx.x & 31) * 8;
    if (col >= M) return;
    const float us = (kPrec == 1) ? scale[b] : 1.0f;
    typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
    u16x8 re, im;
    const float4 *src = reinterpret_cast<const float4 *>(Ab + 2 * ((size_t)(K1 + col) * lda + K0 + k));
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const float4 v = src[h];
        re[2 * h] = tr_round<kPrec>(v.x * us);
        im[2 * h] = tr_round<kPrec>(v.y * us);
        re[2 * h + 1] = tr_round<kPrec>(v.z * us);
        im[2 * h + 1] = tr_round<kPrec>(v.w * us);
    }
    uint16_t *pb = pack + (size_t)b * pstride + 2 * plane + (size_t)col * kLuOuter + k;
    *reinterpret_cast<u16x8 *>(pb) = re;
    *reinterpret_cast<u16x8 *>(pb + plane) = im;
}

// The trailing update on the packed operands: a stage of 32 k is 64 bytes
// per row and plane, loaded as 16-byte vectors (thread = row or column, half
// of the k range) and stored to LDS as is.  Rows and columns >= M of the
// last tiles load zeros and store nothing.  Rows and columns are relative
// to K1 (the update covers A[K1:N, K1:N], M = N - K1).
template <int kPrec>
__global__ __launch_bounds__(256, 2) void k_lu_gemm_trp(float *A, int64_t lda, int64_t a_stride,
                                                     int K1, int M, const uint16_t *pack,
                                                     int64_t pstride, int64_t plane,
                                                     const float *scale) {
    const TrTile tile = trailing_tile();
    const int b = tile.b;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int r0 = tile.x * kTrTile;  // rows / columns relative to K1
    const int c0 = tile.y * kTrTile;
    __shared__ TrLds sm;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    const float us = (kPrec == 1) ? scale[b] : 1.0f;
    const int sr = tid & (kTrTile - 1), sk = (tid >> 7) * 16;
    const bool l_ok = r0 + sr < M, u_ok = c0 + sr < M;
    const uint16_t *pb = pack + (size_t)b * pstride;
    typedef unsigned w4 __attribute__((ext_vector_type(4)));
    const w4 *lre = reinterpret_cast<const w4 *>(pb + (size_t)(r0 + sr) * kLuOuter + sk);
    const w4 *lim = reinterpret_cast<const w4 *>(pb + plane + (size_t)(r0 + sr) * kLuOuter + sk);
    const w4 *ure = reinterpret_cast<const w4 *>(pb + 2 * plane + (size_t)(c0 + sr) * kLuOuter + sk);
    const w4 *uim = reinterpret_cast<const w4 *>(pb + 3 * plane + (size_t)(c0 + sr) * kLuOuter + sk);
    w4 st_[8];
    const w4 z = {0u, 0u, 0u, 0u};
    auto load_stage = [&](int kk) {  // kk: stage start within the outer block
        const int q = kk / 8;            // uint4 index
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            st_[h] = l_ok ? lre[q + h] : z;
            st_[2 + h] = l_ok ? lim[q + h] : z;
            st_[4 + h] = u_ok ? ure[q + h] : z;
            st_[6 + h] = u_ok ? uim[q + h] : z;
        }
    };
    auto store_stage = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            *reinterpret_cast<w4 *>(&sm.Lr[sr][sk + 8 * h]) = st_[h];
            *reinterpret_cast<w4 *>(&sm.Li[sr][sk + 8 * h]) = st_[2 + h];
            *reinterpret_cast<w4 *>(&sm.Ur[sr][sk + 8 * h]) = st_[4 + h];
            *reinterpret_cast<w4 *>(&sm.Ui[sr][sk + 8 * h]) = st_[6 + h];
        }
    };
    f32x4 accr[4][4], acci[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            accr[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            acci[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    using V = typename std::conditional<kPrec == 1, f16x8, bf16x8>::type;
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    load_stage(0);
    for (int kk = 0; kk < kLuOuter; kk += kTrK) {
        __syncthreads();  // the previous stage's operand reads are done
        store_stage();
        __syncthreads();
        if (kk + kTrK < kLuOuter) load_stage(kk + kTrK);  // in flight during the MFMAs
        V lr[4], lm[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            lr[t] = *reinterpret_cast<const V *>(&sm.Lr[wr + 16 * t + li][lk]);
            lm[t] = *reinterpret_cast<const V *>(&sm.Li[wr + 16 * t + li][lk]);
        }
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            const V ur = *reinterpret_cast<const V *>(&sm.Ur[wc + 16 * ti + li][lk]);
            const V ui = *reinterpret_cast<const V *>(&sm.Ui[wc + 16 * ti + li][lk]);
            u32x4 bits = *reinterpret_cast<const u32x4 *>(&ui);
            bits ^= u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};  // -Ui
            const V uin = *reinterpret_cast<const V *>(&bits);
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                if constexpr (kPrec == 1) {
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ur, lr[tj], accr[ti][tj], 0, 0, 0);
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uin, lm[tj], accr[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ui, lr[tj], acci[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ur, lm[tj], acci[ti][tj], 0, 0, 0);
                } else {
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ur, lr[tj], accr[ti][tj], 0, 0, 0);
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(uin, lm[tj], accr[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ui, lr[tj], acci[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ur, lm[tj], acci[ti][tj], 0, 0, 0);
                }
            }
        }
    }
    const float inv_us = 1.0f / us;  // exact: us is a power of two
    tile_epilogue(Ab, lda, K1 + c0 + wc, K1 + r0 + wr, K1 + M, K1 + M, lane, accr, acci, inv_us);
}

// The FP32 trailing update with the same 128 x 128 tiling on
// v_mfma_f32_16x16x4_f32: stages of 16 k staged through LDS planes [k][row |
// col] (row stride 144 floats: the four k rows a wave reads per instruction
// fall in different bank quarters), loaded into registers during the previous
// stage's MFMAs.
constexpr int kTr32K = 16;
constexpr int kTr32S = 144;

struct __attribute__((aligned(16))) TrLds32 {
    float Lr[kTr32K][kTr32S], Li[kTr32K][kTr32S], Ur[kTr32K][kTr32S], Ui[kTr32K][kTr32S];
};

__global__ __launch_bounds__(256, 2) void k_lu_gemm_tr32(float *A, int64_t lda, int64_t a_stride,
                                                         int r_begin, int r_end, int c_begin,
                                                         int c_end, int k_begin, int k_end) {
    const TrTile tile = trailing_tile();
    const int b = tile.b;
    float *Ab = A + (size_t)b * a_stride * 2;
    const int r0 = r_begin + tile.x * kTrTile;
    const int c0 = c_begin + tile.y * kTrTile;
    __shared__ TrLds32 sm;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    // staging of L: thread owns row (tid & 127) and 8 k ((tid >> 7) * 8); of
    // U: 8 lanes read the 16 k of one column as 16-byte pairs, a wave 8
    // columns per load, 4 loads per thread
    const int sr = tid & (kTrTile - 1), sk = (tid >> 7) * 8;
    const int uk = 2 * (tid & 7), uc = tid >> 3;
    const bool lrow_ok = r0 + sr < r_end;
    float2 lv[8];
    float4 uv[4];
    auto load_stage = [&](int k) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            lv[i] = lrow_ok ? *reinterpret_cast<const float2 *>(
                                  Ab + 2 * ((size_t)(k + sk + i) * lda + r0 + sr))
                            : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = c0 + uc + 32 * q;
            uv[q] = c < c_end ? *reinterpret_cast<const float4 *>(
                                    Ab + 2 * ((size_t)c * lda + k + uk))
                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    auto store_stage = [&]() {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            sm.Lr[sk + i][sr] = lv[i].x;
            sm.Li[sk + i][sr] = lv[i].y;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = uc + 32 * q;
            sm.Ur[uk][c] = uv[q].x;
            sm.Ui[uk][c] = uv[q].y;
            sm.Ur[uk + 1][c] = uv[q].z;
            sm.Ui[uk + 1][c] = uv[q].w;
        }
    };
    f32x4 accr[4][4], acci[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            accr[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            acci[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    load_stage(k_begin);
    for (int k = k_begin; k < k_end; k += kTr32K) {
        __syncthreads();  // the previous stage's operand reads are done
        store_stage();
        __syncthreads();
        if (k + kTr32K < k_end) load_stage(k + kTr32K);  // in flight during the MFMAs
#pragma unroll
        for (int ks = 0; ks < kTr32K; ks += 4) {
            const int kk = ks + lk;
            float lr[4], lim[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                lr[t] = sm.Lr[kk][wr + 16 * t + li];
                lim[t] = sm.Li[kk][wr + 16 * t + li];
            }
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) {
                const float ur = sm.Ur[kk][wc + 16 * ti + li];
                const float ui = sm.Ui[kk][wc + 16 * ti + li];
#pragma unroll
                for (int tj = 0; tj < 4; ++tj) {
                    // D^T[col][row]: A operand = U^T (col, k), B operand = L^T (k, row)
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, lr[tj], accr[ti][tj], 0, 0, 0);
                    accr[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui, lim[tj], accr[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui, lr[tj], acci[ti][tj], 0, 0, 0);
                    acci[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur, lim[tj], acci[ti][tj], 0, 0, 0);
                }
            }
        }
    }
    tile_epilogue(Ab, lda, c0 + wc, r0 + wr, c_end, r_end, lane, accr, acci, 1.0f);
}

// Power-of-two scale 2^-e per system with 2^e ~ max |a_ij| (FP16 operand
// range of the trailing update): scale[b] from the max-abs bits in amax[b].
__global__ void k_pow2_scale(const double *amax, float *scale, int B) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const double m = amax[b];
    int e = 0;
    if (m > 0.0 && m < 1e300) (void)frexp(m, &e);
    e = max(-100, min(100, e));
    scale[b] = ldexpf(1.0f, -e);
}

// ---------------------------------------------------------------------------
// Triangular solves with the factors: one 1024-thread block per system,
// x in LDS, sweeps in blocks of kNB rows.  The diagonal block of each sweep
// step is staged in LDS and solved by one wave; the coupling to the other
// rows is a block GEMV over all threads.
//   trans 0: x <- A^{-1} x = U^{-1} L~_{S-1}^{-1} P_{S-1}^T ... L~_0^{-1} P_0^T x
//   trans 2: x <- A^{-H} x = P_0 L~_0^{-H} ... P_{S-1} L~_{S-1}^{-H} U^{-H} x
// where s runs over the outer blocks of kLuOuter columns, P_s^T applies the
// exchanges of block s in order and L~_s is the unit lower block column of
// block s in the row order at the end of that block.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void stage_diag(const float *Ab, int64_t lda, int kb, c32 (*D)[kNB + 1]) {
    const int i = threadIdx.x % kNB, j = threadIdx.x / kNB;  // 1024 threads = 32 x 32
    D[i][j] = ld(Ab, (size_t)(kb + j) * lda + kb + i);
}

// xs[kb + c] -= sum_r conj(A[r][kb + c]) xs[r] over rows r in [r0, r1)
// (32 outputs, all threads; partial sums in part).
__device__ __forceinline__ void block_gemv_h(const float *Ab, int64_t lda, int kb, int r0, int r1,
                                             c32 *xs, c32 (*part)[kNB + 1]) {
    const int c = threadIdx.x / kNB, chunk = threadIdx.x % kNB;
    c32 acc{0.0f, 0.0f}, acc2{0.0f, 0.0f};
    const size_t cc = (size_t)(kb + c) * lda;
    int r = r0 + chunk;
    // eight loads in flight per round trip
    for (; r + 7 * kNB < r1; r += 8 * kNB) {
        c32 a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = ld(Ab, cc + r + u * kNB);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const c32 x = xs[r + u * kNB];
            c32 &s = (u & 1) ? acc2 : acc;
            s.re += a[u].re * x.re + a[u].im * x.im;
            s.im += a[u].re * x.im - a[u].im * x.re;
        }
    }
    for (; r < r1; r += kNB) {
        const c32 a = ld(Ab, cc + r);
        const c32 x = xs[r];
        acc.re += a.re * x.re + a.im * x.im;
        acc.im += a.re * x.im - a.im * x.re;
    }
    acc.re += acc2.re;
    acc.im += acc2.im;
    part[chunk][c] = acc;
    __syncthreads();
    if (threadIdx.x < kNB) {
        c32 s{0.0f, 0.0f};
#pragma unroll 8
        for (int q = 0; q < kNB; ++q) {
            s.re += part[q][threadIdx.x].re;
            s.im += part[q][threadIdx.x].im;
        }
        xs[kb + threadIdx.x].re -= s.re;
        xs[kb + threadIdx.x].im -= s.im;
    }
}

// Applies the composed row exchanges of panel steps [q0, q1) to xs, in
// order (position trow[m] takes the old entry tsrc[m]).  Run by one wave:
// the records are loaded at once, and each step's reads precede its writes
// in the wave's own order.
__device__ __forceinline__ void apply_gather_lists(const int32_t *glb, int q0, int q1, c32 *xs) {
    const int lane = threadIdx.x & 63;
    constexpr int kMaxSteps = kLuOuter / kNB;
    int cnt[kMaxSteps], trow[kMaxSteps], tsrc[kMaxSteps];
#pragma unroll
    for (int i = 0; i < kMaxSteps; ++i) {
        cnt[i] = 0;
        if (q0 + i < q1) {
            const int32_t *g = glb + (size_t)(q0 + i) * kLuGather;
            cnt[i] = g[0];
            trow[i] = g[1 + lane];
            tsrc[i] = g[1 + 2 * kNB + lane];
        }
    }
#pragma unroll
    for (int i = 0; i < kMaxSteps; ++i) {
        const bool on = lane < cnt[i];
        const c32 t = on ? xs[tsrc[i]] : c32{0.0f, 0.0f};
        if (on) xs[trow[i]] = t;
    }
}

// trans 0, forward: at the first step of each outer block the block's row
// exchanges (its composed gather lists), then per kNB-row step the
// unit-lower diagonal solve (one wave, lane = row, values broadcast with
// v_readlane) and the update of the rows below; backward: the upper
// diagonal solve and the update of the rows above.  Each step issues the
// loads of its coupling block (one row per thread) and of the next step's
// diagonal block before it waits, so they overlap the diagonal solve.
__global__ __launch_bounds__(1024) void k_lu_solve(const float *A, int64_t lda, int64_t a_stride,
                                                   int N, const int32_t *ipiv, const int32_t *gl,
                                                   int64_t gl_stride, float *X, int64_t ldx,
                                                   int trans, int vpm) {
    extern __shared__ c32 xs[];
    __shared__ c32 D[kNB][kNB + 1];
    __shared__ c32 part[kNB][kNB + 1];
    __shared__ int pv_s[kLuOuter];
    const int b = blockIdx.x;  // vector; its matrix is b / vpm
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    const int32_t *pv = ipiv + (size_t)(b / vpm) * N;
    float *xb = X + (size_t)b * ldx * 2;
    for (int r = tid; r < N; r += 1024) xs[r] = ld(xb, r);
    if (trans == 0) {
        const int32_t *glb = gl + (size_t)(b / vpm) * gl_stride;
        const int di = tid % kNB, dj = tid / kNB;  // the thread's diagonal-block entry
        c32 dnext = ld(Ab, (size_t)dj * lda + di);
        __syncthreads();  // xs loaded
        for (int kb = 0; kb < N; kb += kNB) {
            const c32 dcur = dnext;
            D[di][dj] = dcur;
            dnext = kb + kNB < N ? ld(Ab, (size_t)(kb + kNB + dj) * lda + kb + kNB + di) : dcur;
            const int r = kb + kNB + tid;
            c32 lr[kNB];
            if (r < N) {
#pragma unroll
                for (int j = 0; j < kNB; ++j) lr[j] = ld(Ab, (size_t)(kb + j) * lda + r);
            }
            if (kb % kLuOuter == 0 && tid < 64)
                apply_gather_lists(glb, kb / kNB, min(N, kb + kLuOuter) / kNB, xs);
            __syncthreads();
            if (tid < 64) {
                c32 x = lane < kNB ? xs[kb + lane] : c32{0.0f, 0.0f};
#pragma unroll
                for (int j = 0; j < kNB - 1; ++j) {
                    const float xr = lane_value(x.re, j);
                    const float xi = lane_value(x.im, j);
                    if (lane > j && lane < kNB) {
                        const c32 l = D[lane][j];
                        x.re -= l.re * xr - l.im * xi;
                        x.im -= l.re * xi + l.im * xr;
                    }
                }
                if (lane < kNB) xs[kb + lane] = x;
            }
            __syncthreads();
            if (r < N) {
                c32 acc = xs[r];
#pragma unroll
                for (int j = 0; j < kNB; ++j) {
                    const c32 l = lr[j];
                    const c32 xj = xs[kb + j];
                    acc.re -= l.re * xj.re - l.im * xj.im;
                    acc.im -= l.re * xj.im + l.im * xj.re;
                }
                xs[r] = acc;
            }
            for (int r2 = r + 1024; r2 < N; r2 += 1024) {
                c32 acc = xs[r2];
#pragma unroll 8
                for (int j = 0; j < kNB; ++j) {
                    const c32 l = ld(Ab, (size_t)(kb + j) * lda + r2);
                    const c32 xj = xs[kb + j];
                    acc.re -= l.re * xj.re - l.im * xj.im;
                    acc.im -= l.re * xj.im + l.im * xj.re;
                }
                xs[r2] = acc;
            }
            __syncthreads();
        }
        // dnext holds the entry of the last diagonal block
        for (int kb = N - kNB; kb >= 0; kb -= kNB) {
            const c32 dcur = dnext;
            D[di][dj] = dcur;
            if (kb > 0) dnext = ld(Ab, (size_t)(kb - kNB + dj) * lda + kb - kNB + di);
            const int r = tid;
            c32 ur[kNB];
            if (r < kb) {
#pragma unroll
                for (int j = 0; j < kNB; ++j) ur[j] = ld(Ab, (size_t)(kb + j) * lda + r);
            }
            __syncthreads();
            if (tid < 64) {
                c32 x = lane < kNB ? xs[kb + lane] : c32{0.0f, 0.0f};
#pragma unroll
                for (int j = kNB - 1; j >= 0; --j) {
                    if (lane == j) x = cmul(x, cinv(D[j][j]));
                    const float xr = lane_value(x.re, j);
                    const float xi = lane_value(x.im, j);
                    if (lane < j) {
                        const c32 u = D[lane][j];
                        x.re -= u.re * xr - u.im * xi;
                        x.im -= u.re * xi + u.im * xr;
                    }
                }
                if (lane < kNB) xs[kb + lane] = x;
            }
            __syncthreads();
            if (r < kb) {
                c32 acc = xs[r];
#pragma unroll
                for (int j = 0; j < kNB; ++j) {
                    const c32 u = ur[j];
                    const c32 xj = xs[kb + j];
                    acc.re -= u.re * xj.re - u.im * xj.im;
                    acc.im -= u.re * xj.im + u.im * xj.re;
                }
                xs[r] = acc;
            }
            for (int r2 = r + 1024; r2 < kb; r2 += 1024) {
                c32 acc = xs[r2];
#pragma unroll 8
                for (int j = 0; j < kNB; ++j) {
                    const c32 u = ld(Ab, (size_t)(kb + j) * lda + r2);
                    const c32 xj = xs[kb + j];
                    acc.re -= u.re * xj.re - u.im * xj.im;
                    acc.im -= u.re * xj.im + u.im * xj.re;
                }
                xs[r2] = acc;
            }
            __syncthreads();
        }
    } else {
        // U^H y = x, forward over the blocks: couple to the solved rows above,
        // then the lower triangular (U11)^H.
        for (int kb = 0; kb < N; kb += kNB) {
            __syncthreads();
            if (kb > 0) block_gemv_h(Ab, lda, kb, 0, kb, xs, part);
            stage_diag(Ab, lda, kb, D);
            __syncthreads();
            if (tid < 64) {
                c32 x = lane < kNB ? xs[kb + lane] : c32{0.0f, 0.0f};
                for (int j = 0; j < kNB; ++j) {
                    if (lane == j) x = cmul(x, cinv(cconj(D[j][j])));
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane > j && lane < kNB) {
                        const c32 u = cconj(D[j][lane]);
                        x.re -= u.re * xr - u.im * xi;
                        x.im -= u.re * xi + u.im * xr;
                    }
                }
                if (lane < kNB) xs[kb + lane] = x;
            }
        }
        // L~_s^{-H} then P_s, s = S-1 .. 0 (outer blocks, their kNB-row
        // sub-blocks from the last).
        for (int kb = N - kNB; kb >= 0; kb -= kNB) {
            __syncthreads();
            if (kb + kNB < N) block_gemv_h(Ab, lda, kb, kb + kNB, N, xs, part);
            const int xw = min(kLuOuter, N - kb);
            const bool xstart = (kb % kLuOuter) == 0;
            if (xstart && tid < xw) pv_s[tid] = pv[kb + tid];
            stage_diag(Ab, lda, kb, D);
            __syncthreads();
            if (tid < 64) {
                c32 x = lane < kNB ? xs[kb + lane] : c32{0.0f, 0.0f};
                for (int j = kNB - 1; j >= 0; --j) {
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane < j) {
                        const c32 l = cconj(D[j][lane]);
                        x.re -= l.re * xr - l.im * xi;
                        x.im -= l.re * xi + l.im * xr;
                    }
                }
                if (lane < kNB) xs[kb + lane] = x;
            }
            __syncthreads();
            if (xstart && tid == 0)
                for (int j = xw - 1; j >= 0; --j) {
                    const int p = pv_s[j];
                    if (p != kb + j) {
                        const c32 t = xs[kb + j];
                        xs[kb + j] = xs[p];
                        xs[p] = t;
                    }
                }
        }
        __syncthreads();
    }
    for (int r = tid; r < N; r += 1024) st(xb, r, xs[r]);
}

// ---------------------------------------------------------------------------
// Triangular solves on many workgroups per vector (orders <= kTrsvMaxOrder,
// with a TrsvSync).  One CU streams the factors of one system at a few tens of GB/s,
// so the one-workgroup solve above is bound by that CU; here the order is cut
// into 64-row panels and each panel is a 256-thread workgroup that streams its
// own rows of the factor, folds in the values of the panels it depends on as
// they are published, applies the inverse of its 64 x 64 diagonal block (one
// matrix-vector product on all threads: the inverses are built once per
// factorization by k_lu_diag_inverse, since a serial 64-step substitution on
// one wave is the longest link of the chain) and publishes its values.  A value
// is published as two 64-bit words (real or imaginary bits | epoch << 32),
// stored and polled with agent-scope atomics: a reader sees a whole word or
// none of it, so the tag says the value is there and no flag, fence or store
// wait sits between producer and consumer.  A workgroup takes its panel from
// a ticket counter in arrival order and waits only on panels with smaller
// tickets, which belong to workgroups already running, so every wait ends
// without co-residency of the grid; each wait is also bounded (kTrsvSpin
// polls): past it the panel goes on with NaN values and sets the context's
// error word.
// Rows are in the final order of the factorization (all exchanges applied).
// The L columns of outer block s are stored in the row order at the end of
// block s; k_lu_final_l copies L once per factorization into final row order
// (lf), so every coupling tile is read contiguously.
//   mode 0: y = L^{-1} P b      (forward,  b = X read through perm)
//   mode 1: X = U^{-1} y        (backward)
//   mode 2: y = U^{-H} b        (forward,  b = X)
//   mode 3: X = P^T L^{-H} y    (backward, X written through perm)
// Forward panels publish into sync.y, backward panels into sync.z ([V][N][2]
// tagged words).
// ---------------------------------------------------------------------------
constexpr int kTrsvRows = 64;
constexpr uint32_t kTrsvSpin = 1u << 20;
static_assert(kLuOuter % kTrsvRows == 0 && kLuPad % kTrsvRows == 0, "panels inside outer blocks");

__global__ void k_trsv_init(uint32_t *ctx, uint64_t *y, uint64_t *z, size_t n_words) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < 4) ctx[i] = i == 2 ? 1u : 0u;  // ticket, finished, epoch 1, error
    if (i < n_words) {
        y[i] = 0ull;  // epoch 0: never current
        z[i] = 0ull;
    }
}

// Row maps of one system per block: the composed exchanges of every panel
// step are replayed on an index vector in LDS by wave 0 (the records of an
// outer block's steps loaded at once), once for perm (all steps) and once per
// outer block s for pi[s].  LDS: 2 N ints.
__global__ __launch_bounds__(1024) void k_lu_solve_tables(const int32_t *gl, int64_t gl_stride,
                                                          int N, int32_t *perm, int32_t *pi) {
    extern __shared__ int32_t tsm[];
    int32_t *ord = tsm, *inv = tsm + N;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    constexpr int kSteps = kLuOuter / kNB;
    const int Q = N / kNB, S = (N + kLuOuter - 1) / kLuOuter;
    const int32_t *glb = gl + (size_t)b * gl_stride;
    int32_t *permb = perm + (size_t)b * N;
    // position trow takes the entry at tsrc, steps in order; each step's
    // reads precede its writes in the wave's order
    auto replay = [&](int q0, int q1) {
        for (int q = q0; q < q1; q += kSteps) {
            int cnt[kSteps], trow[kSteps], tsrc[kSteps];
#pragma unroll
            for (int i = 0; i < kSteps; ++i) {
                cnt[i] = 0;
                trow[i] = tsrc[i] = 0;
                if (q + i < q1) {
                    const int32_t *g = glb + (size_t)(q + i) * kLuGather;
                    cnt[i] = g[0];
                    trow[i] = g[1 + lane];
                    tsrc[i] = g[1 + 2 * kNB + lane];
                }
            }
#pragma unroll
            for (int i = 0; i < kSteps; ++i) {
                const bool on = lane < cnt[i];
                const int t = on ? ord[tsrc[i]] : 0;
                if (on) ord[trow[i]] = t;
            }
        }
    };
    for (int r = tid; r < N; r += 1024) ord[r] = r;
    __syncthreads();
    if (tid < 64) replay(0, Q);
    __syncthreads();
    for (int r = tid; r < N; r += 1024) permb[r] = ord[r];
    __syncthreads();  // perm (global) is read back below by other threads of the block
    for (int r = tid; r < N; r += 1024) ord[r] = r;
    __syncthreads();
    for (int s = 0; s < S; ++s) {
        if (tid < 64) replay(s * kSteps, min(Q, (s + 1) * kSteps));
        __syncthreads();
        for (int p = tid; p < N; p += 1024) inv[ord[p]] = p;
        __syncthreads();
        int32_t *m = pi + ((size_t)b * S + s) * N;
        for (int r = tid; r < N; r += 1024) m[r] = inv[permb[r]];
        __syncthreads();
    }
}

// Inverses of the 64 x 64 diagonal blocks of the factors: dinv[b][p][0] =
// L_pp^{-1} (unit lower), dinv[b][p][1] = U_pp^{-1}, element (i, j) at
// i * 64 + j.  One wave per block and factor (blockIdx.z); lane c builds
// column c by substitution, the column in registers (both loops unrolled,
// so every index is a constant), the block in LDS (read as broadcasts).  The
// loops run over the same k on every lane: entries outside the column's
// triangle come out as exact zeros (sums of products with zeros).
template <bool kUpper>
__device__ __forceinline__ void diag_inverse_column(const c32 (*T)[kTrsvRows], int c,
                                                    c32 (&col)[kTrsvRows]) {
    if (!kUpper) {
#pragma unroll
        for (int i = 0; i < kTrsvRows; ++i) {
            c32 acc{i == c ? 1.0f : 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < i; ++k) {
                const c32 l = T[i][k];
                acc.re -= l.re * col[k].re - l.im * col[k].im;
                acc.im -= l.re * col[k].im + l.im * col[k].re;
            }
            col[i] = acc;
        }
    } else {
#pragma unroll
        for (int i = kTrsvRows - 1; i >= 0; --i) {
            c32 acc{i == c ? 1.0f : 0.0f, 0.0f};
#pragma unroll
            for (int k = i + 1; k < kTrsvRows; ++k) {
                const c32 u = T[i][k];
                acc.re -= u.re * col[k].re - u.im * col[k].im;
                acc.im -= u.re * col[k].im + u.im * col[k].re;
            }
            col[i] = cmul(acc, cinv(T[i][i]));
        }
    }
}

__global__ __launch_bounds__(64) void k_lu_diag_inverse(const float *A, int64_t lda,
                                                        int64_t a_stride, int N, float *dinv) {
    __shared__ c32 T[kTrsvRows][kTrsvRows];
    const int p = blockIdx.x, b = blockIdx.y, upper = blockIdx.z, c = threadIdx.x;
    const int P = N / kTrsvRows, R0 = p * kTrsvRows;
    const float *Ab = A + (size_t)b * a_stride * 2;
    for (int j = 0; j < kTrsvRows; ++j) T[c][j] = ld(Ab, (size_t)(R0 + j) * lda + R0 + c);
    __syncthreads();
    c32 col[kTrsvRows];
    if (upper)
        diag_inverse_column<true>(T, c, col);
    else
        diag_inverse_column<false>(T, c, col);
    float *out = dinv + (((size_t)b * P + p) * 2 + upper) * kTrsvRows * kTrsvRows * 2;
#pragma unroll
    for (int i = 0; i < kTrsvRows; ++i) st(out, (size_t)i * kTrsvRows + c, col[i]);
}

// L below the diagonal in final row order: lf[b][c][r] = L[pi[s(c)][r]][c]
// for r > c, one block per column (its stored column read with a gather
// inside the column, the copy written contiguously).
__global__ __launch_bounds__(256) void k_lu_final_l(const float *A, int64_t lda, int64_t a_stride,
                                                    int N, const int32_t *pi, float *lf) {
    const int c = blockIdx.x, b = blockIdx.y, S = (N + kLuOuter - 1) / kLuOuter;
    const float *col = A + ((size_t)b * a_stride + (size_t)c * lda) * 2;
    const int32_t *m = pi + ((size_t)b * S + c / kLuOuter) * N;
    float *out = lf + ((size_t)b * a_stride + (size_t)c * lda) * 2;
    for (int r = c + 1 + (int)threadIdx.x; r < N; r += 256) st(out, r, ld(col, m[r]));
}

__device__ __forceinline__ uint64_t tag_word(float v, uint32_t ep) {
    return ((uint64_t)ep << 32) | (uint64_t)__float_as_uint(v);
}
__device__ __forceinline__ float word_value(uint64_t w) { return __uint_as_float((uint32_t)w); }
__device__ __forceinline__ uint32_t word_epoch(uint64_t w) { return (uint32_t)(w >> 32); }

template <int kMode>
__global__ __launch_bounds__(256) void k_trsv(const float *A, const float *lf, int64_t lda,
                                              int64_t a_stride, int N, const int32_t *perm,
                                              const float *dinv, float *X, int64_t ldx, int vpm,
                                              TrsvSync sync) {
    constexpr bool kFwd = kMode == 0 || kMode == 2;
    // modes 0, 1: lanes = own rows, waves = column quarters of a coupling
    // panel; modes 2, 3 (conjugate transposes, columns of the stored factor):
    // lanes = coupling rows, waves = quarters of the own rows
    constexpr bool kRowLanes = kMode <= 1;
    __shared__ c32 D[kTrsvRows][kTrsvRows + 1];
    __shared__ c32 xs[2][kTrsvRows];
    __shared__ c32 vs[kTrsvRows];
    __shared__ c32 part[4][kTrsvRows];
    __shared__ uint32_t role_s, ep_s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int P = N / kTrsvRows;
    if (tid == 0) {
        role_s = atomicAdd(&sync.ctx[0], 1u);
        ep_s = __hip_atomic_load(&sync.ctx[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int role = (int)role_s;
    const uint32_t ep = ep_s;
    // tickets panel-major over the vectors: every vector's chain advances at
    // once (vector-major tickets left most of a large grid waiting for the
    // first vectors' chains); a panel still depends only on smaller tickets
    const int V = (int)gridDim.x / P;
    const int q = role / V, v = role % V;
    const int p = kFwd ? q : P - 1 - q;  // panels taken in dependency order
    const int R0 = p * kTrsvRows;
    const int b = v / vpm;
    // the factor the coupling tiles come from: L in final row order (modes
    // 0, 3) or U
    const float *Ab = (kMode == 0 || kMode == 3 ? lf : A) + (size_t)b * a_stride * 2;
    const int32_t *permb = perm + (size_t)b * N;
    const uint64_t *yv = sync.y + (size_t)v * N * 2;
    uint64_t *pub = (kFwd ? sync.y : sync.z) + (size_t)v * N * 2;
    float *xv = X + (size_t)v * ldx * 2;
    // D[i][j] = entry (i, j) of the inverse of the own diagonal block of the
    // factor solved with: L^{-1}, U^{-1}, (U^H)^{-1} = (U^{-1})^H, (L^H)^{-1}
    const float *Dv = dinv + (((size_t)b * P + p) * 2 + (kMode == 1 || kMode == 2 ? 1 : 0)) *
                                 kTrsvRows * kTrsvRows * 2;
    for (int e = tid; e < kTrsvRows * kTrsvRows; e += 256) {
        if (kRowLanes) {
            const int j = e & 63, i = e >> 6;
            D[i][j] = ld(Dv, (size_t)i * kTrsvRows + j);
        } else {
            const int i = e & 63, j = e >> 6;
            D[i][j] = cconj(ld(Dv, (size_t)j * kTrsvRows + i));
        }
    }
    c32 rhs{0.0f, 0.0f};
    int out_row = 0;  // mode 3: the row of X the lane's value goes to
    if (w == 0) {
        const int r = R0 + lane;
        if (kMode == 3) out_row = permb[r];
        // the backward passes read the forward pass's words (earlier launch)
        rhs = kMode == 0   ? ld(xv, permb[r])
              : kMode == 2 ? ld(xv, r)
                           : c32{word_value(yv[2 * r]), word_value(yv[2 * r + 1])};
    }
    __syncthreads();
    const int n_dep = kFwd ? p : P - 1 - p;
    auto dep = [&](int k) { return kFwd ? k : P - 1 - k; };
    // 16 factor values of coupling panel J for this thread: entries (own
    // row, J column) of L or U (modes 0, 1), entries (J row, own column) of
    // U or L (modes 2, 3; conjugated in the fold)
    auto load_tile = [&](int J, c32 (&t)[16]) {
        if (kRowLanes) {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                t[j] = ld(Ab, (size_t)(J * kTrsvRows + w * 16 + j) * lda + R0 + lane);
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                t[j] = ld(Ab, (size_t)(R0 + w * 16 + j) * lda + J * kTrsvRows + lane);
        }
    };
    c32 acc[kRowLanes ? 1 : 16];
#pragma unroll
    for (int j = 0; j < (kRowLanes ? 1 : 16); ++j) acc[j] = c32{0.0f, 0.0f};
    bool ok = true;
    // waits until panel J's values carry this epoch (wave 0, lane = row of
    // J, polling its two words; relaxed coherent loads, no cache
    // maintenance) and stages them in xs[buf]
    auto fetch = [&](int J, int buf) {
        if (w == 0) {
            const uint64_t *src = pub + 2 * ((size_t)J * kTrsvRows + lane);
            uint64_t re = 0ull, im = 0ull;
            for (uint32_t i = 0;; ++i) {
                re = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                im = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__all(word_epoch(re) == ep && word_epoch(im) == ep)) break;
                if (!ok || i + 1 >= kTrsvSpin) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            xs[buf][lane] = c32{word_value(re), word_value(im)};
        }
        __syncthreads();
    };
    auto fold = [&](const c32 (&t)[16], int buf) {
        if (kRowLanes) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const c32 x = xs[buf][w * 16 + j];
                acc[0].re += t[j].re * x.re - t[j].im * x.im;
                acc[0].im += t[j].re * x.im + t[j].im * x.re;
            }
        } else {
            const c32 x = xs[buf][lane];
#pragma unroll
            for (int j = 0; j < 16; ++j) {  // conj(t) x
                acc[j].re += t[j].re * x.re + t[j].im * x.im;
                acc[j].im += t[j].re * x.im - t[j].im * x.re;
            }
        }
    };
    c32 ta[16], tb[16];
    if (n_dep > 0) load_tile(dep(0), ta);
    for (int k = 0; k < n_dep; k += 2) {
        if (k + 1 < n_dep) load_tile(dep(k + 1), tb);
        fetch(dep(k), 0);
        fold(ta, 0);
        if (k + 1 < n_dep) {
            if (k + 2 < n_dep) load_tile(dep(k + 2), ta);
            fetch(dep(k + 1), 1);
            fold(tb, 1);
        }
    }
    // coupling sums of the own rows into part[0][row]
    if (kRowLanes) {
        part[w][lane] = acc[0];
        __syncthreads();
        if (w == 0) {
            c32 s = part[0][lane];
#pragma unroll
            for (int u = 1; u < 4; ++u) {
                s.re += part[u][lane].re;
                s.im += part[u][lane].im;
            }
            part[0][lane] = s;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            float re = acc[j].re, im = acc[j].im;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                re += __shfl_xor(re, off);
                im += __shfl_xor(im, off);
            }
            acc[j] = c32{re, im};
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) part[0][w * 16 + j] = acc[j];
        }
    }
    __syncthreads();
    if (w == 0) vs[lane] = c32{rhs.re - part[0][lane].re, rhs.im - part[0][lane].im};
    __syncthreads();
    {
        // own values = D vs: thread (row lane, column quarter w)
        c32 s{0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const c32 d = D[lane][w * 16 + j], u = vs[w * 16 + j];
            s.re += d.re * u.re - d.im * u.im;
            s.im += d.re * u.im + d.im * u.re;
        }
        part[w][lane] = s;
    }
    __syncthreads();
    if (w == 0) {
        c32 x = part[0][lane];
#pragma unroll
        for (int u = 1; u < 4; ++u) {
            x.re += part[u][lane].re;
            x.im += part[u][lane].im;
        }
        if (!ok) x = c32{__builtin_nanf(""), __builtin_nanf("")};
        if (kMode == 1) st(xv, R0 + lane, x);
        if (kMode == 3) st(xv, out_row, x);
        uint64_t *dst = pub + 2 * (size_t)(R0 + lane);
        __hip_atomic_store(dst, tag_word(x.re, ep), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(dst + 1, tag_word(x.im, ep), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (tid == 0) {
        if (!ok) atomicOr(&sync.ctx[3], 1u);
        // the last workgroup resets the ticket counter and opens the next epoch
        if (atomicAdd(&sync.ctx[1], 1u) == gridDim.x - 1) {
            __hip_atomic_store(&sync.ctx[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sync.ctx[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sync.ctx[2], ep + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Stepped solves for orders above 2048 (x does not fit one block's LDS
// budget, and one CU per vector cannot stream the factors fast enough): per
// outer block of kLuOuter rows, one in-block kernel (one 256-thread block
// per vector: the block's exchanges, its kNB-row sub-block sweeps with the
// coupling inside the outer block, the diagonal solves) and one coupling
// kernel that carries the block's contribution to, or gathers the
// contribution of, the rest of the vector on many workgroups.
//   A x = b:   per block s (forward)  P_s^T, L~_s in-block, rows below -= L x_s;
//              per block s (backward) U_ss in-block, rows above -= U x_s.
//   A^H x = b: per block s (forward)  x_s -= (U above)^H x, U_ss^H in-block;
//              per block s (backward) x_s -= (L~ below)^H x, L~_s^H in-block,
//              then P_s.
// ---------------------------------------------------------------------------

// Applies the composed exchanges of panel steps [q0, q1) of one system to x
// in global memory: forward (position trow takes the old entry tsrc, steps
// in order) or inverse (position tsrc takes the entry at trow, steps in
// reverse order).  Every thread of the block calls it; wave 0 moves the
// entries, one step per barrier (a step touches at most 64 positions, and
// all its loads complete before its stores: the stores consume the loads).
__device__ void gather_steps_global(float *xb, const int32_t *glb, int q0, int q1,
                                    bool inverse) {
    const int lane = threadIdx.x & 63;
    for (int i = 0; i < q1 - q0; ++i) {
        const int q = inverse ? q1 - 1 - i : q0 + i;
        if (threadIdx.x < 64) {
            const int32_t *g = glb + (size_t)q * kLuGather;
            const int cnt = g[0];
            if (lane < cnt) {
                const int trow = g[1 + lane], tsrc = g[1 + 2 * kNB + lane];
                const c32 t = ld(xb, inverse ? trow : tsrc);
                st(xb, inverse ? tsrc : trow, t);
            }
        }
        __syncthreads();
    }
}

// In-block sweep of outer block [K0, K0 + w) for vector blockIdx.x.
// mode 0: P_s^T then the unit lower L~_ss (forward); 1: the upper U_ss with
// its diagonal (backward); 2: U_ss^H (forward); 3: L~_ss^H (backward) then
// P_s.  Coupling to rows outside the block is done by k_solve_upd_*.
__global__ __launch_bounds__(256) void k_solve_block(const float *A, int64_t lda,
                                                     int64_t a_stride, int N,
                                                     const int32_t *gl, int64_t gl_stride,
                                                     float *X, int64_t ldx, int K0, int mode,
                                                     int vpm) {
    __shared__ c32 xs[kLuOuter];
    __shared__ c32 D[kNB][kNB + 1];
    __shared__ c32 part[8][kNB + 1];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63;
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    float *xb = X + (size_t)b * ldx * 2;
    const int32_t *glb = gl + (size_t)(b / vpm) * gl_stride;
    const int w = min(kLuOuter, N - K0);
    const int q0 = K0 / kNB, q1 = (K0 + w) / kNB;
    if (mode == 0) gather_steps_global(xb, glb, q0, q1, false);
    if (tid < w) xs[tid] = ld(xb, K0 + tid);
    const bool fwd = (mode == 0 || mode == 2);
    const int nsub = w / kNB;
    for (int s = 0; s < nsub; ++s) {
        const int o = fwd ? s * kNB : (nsub - 1 - s) * kNB;  // sub-block offset in the block
        const int c0 = fwd ? 0 : o + kNB, c1 = fwd ? o : w;  // solved entries it couples to
        __syncthreads();  // xs complete, D and part free
        for (int e = tid; e < kNB * kNB; e += 256) {
            const int i = e % kNB, j = e / kNB;
            D[i][j] = ld(Ab, (size_t)(K0 + o + j) * lda + K0 + o + i);
        }
        if (c1 > c0) {
            c32 acc{0.0f, 0.0f};
            if (mode <= 1) {
                // sum_c A(K0+o+i, K0+c) x[c]: lanes along rows i, 8 column chunks
                const int i = tid % kNB, ch = tid / kNB;
                for (int c = c0 + ch; c < c1; c += 8) {
                    const c32 a = ld(Ab, (size_t)(K0 + c) * lda + K0 + o + i);
                    const c32 xv = xs[c];
                    acc.re += a.re * xv.re - a.im * xv.im;
                    acc.im += a.re * xv.im + a.im * xv.re;
                }
                part[ch][i] = acc;
            } else {
                // sum_r conj(A(K0+r, K0+o+j)) x[r]: 8 lanes along rows r per column j
                const int ch = tid % 8, j = tid / 8;
                for (int r = c0 + ch; r < c1; r += 8) {
                    const c32 a = ld(Ab, (size_t)(K0 + o + j) * lda + K0 + r);
                    const c32 xv = xs[r];
                    acc.re += a.re * xv.re + a.im * xv.im;
                    acc.im += a.re * xv.im - a.im * xv.re;
                }
                part[ch][j] = acc;
            }
        }
        __syncthreads();
        if (tid < 64) {
            c32 x = lane < kNB ? xs[o + lane] : c32{0.0f, 0.0f};
            if (c1 > c0 && lane < kNB) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    x.re -= part[q][lane].re;
                    x.im -= part[q][lane].im;
                }
            }
            if (mode == 0) {
                for (int j = 0; j < kNB - 1; ++j) {
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane > j && lane < kNB) {
                        const c32 l = D[lane][j];
                        x.re -= l.re * xr - l.im * xi;
                        x.im -= l.re * xi + l.im * xr;
                    }
                }
            } else if (mode == 1) {
                for (int j = kNB - 1; j >= 0; --j) {
                    if (lane == j) x = cmul(x, cinv(D[j][j]));
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane < j) {
                        const c32 u = D[lane][j];
                        x.re -= u.re * xr - u.im * xi;
                        x.im -= u.re * xi + u.im * xr;
                    }
                }
            } else if (mode == 2) {
                for (int j = 0; j < kNB; ++j) {
                    if (lane == j) x = cmul(x, cinv(cconj(D[j][j])));
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane > j && lane < kNB) {
                        const c32 u = cconj(D[j][lane]);
                        x.re -= u.re * xr - u.im * xi;
                        x.im -= u.re * xi + u.im * xr;
                    }
                }
            } else {
                for (int j = kNB - 1; j >= 0; --j) {
                    const float xr = __shfl(x.re, j), xi = __shfl(x.im, j);
                    if (lane < j) {
                        const c32 l = cconj(D[j][lane]);
                        x.re -= l.re * xr - l.im * xi;
                        x.im -= l.re * xi + l.im * xr;
                    }
                }
            }
            if (lane < kNB) xs[o + lane] = x;
        }
    }
    __syncthreads();
    if (tid < w) st(xb, K0 + tid, xs[tid]);
    if (mode == 3) {
        __syncthreads();  // the block's entries are in x before the exchanges move them
        gather_steps_global(xb, glb, q0, q1, true);
    }
}

// x[r] -= sum_{c in [K0, K1)} A(r, c) x[c] for rows r in [r0, r1), one
// thread per row (loads coalesced along each column), grid.y = vector.
__global__ __launch_bounds__(256) void k_solve_upd_n(const float *A, int64_t lda,
                                                     int64_t a_stride, float *X, int64_t ldx,
                                                     int K0, int K1, int r0, int r1, int vpm) {
    const int b = blockIdx.y;
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    float *xb = X + (size_t)b * ldx * 2;
    __shared__ c32 xk[kLuOuter];
    const int w = K1 - K0;
    if (threadIdx.x < w) xk[threadIdx.x] = ld(xb, K0 + threadIdx.x);
    __syncthreads();
    const int r = r0 + blockIdx.x * 256 + threadIdx.x;
    if (r >= r1) return;
    c32 acc{0.0f, 0.0f};
#pragma unroll 8
    for (int c = 0; c < w; ++c) {
        const c32 a = ld(Ab, (size_t)(K0 + c) * lda + r);
        const c32 xj = xk[c];
        acc.re += a.re * xj.re - a.im * xj.im;
        acc.im += a.re * xj.im + a.im * xj.re;
    }
    c32 v = ld(xb, r);
    v.re -= acc.re;
    v.im -= acc.im;
    st(xb, r, v);
}

// x[c] -= sum_{r in [r0, r1)} conj(A(r, c)) x[r] for c in [K0, K1), one wave
// per column (lanes along the column), grid.y = vector.
__global__ __launch_bounds__(256) void k_solve_upd_h(const float *A, int64_t lda,
                                                     int64_t a_stride, float *X, int64_t ldx,
                                                     int K0, int K1, int r0, int r1, int vpm) {
    const int b = blockIdx.y;
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    float *xb = X + (size_t)b * ldx * 2;
    const int c = K0 + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= K1) return;
    c32 acc{0.0f, 0.0f};
#pragma unroll 4
    for (int r = r0 + lane; r < r1; r += 64) {
        const c32 a = ld(Ab, (size_t)c * lda + r);
        const c32 xv = ld(xb, r);
        acc.re += a.re * xv.re + a.im * xv.im;
        acc.im += a.re * xv.im - a.im * xv.re;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc.re += __shfl_xor(acc.re, off);
        acc.im += __shfl_xor(acc.im, off);
    }
    if (lane == 0) {
        c32 v = ld(xb, c);
        v.re -= acc.re;
        v.im -= acc.im;
        st(xb, c, v);
    }
}

// ---------------------------------------------------------------------------
// Reciprocal condition estimate of every factored system of a batch, as
// cgecon does it: ||A^{-1}||_1 by Hager's method with Higham's refinements
// (LAPACK clacn2; one state machine per system, driven by batched solves),
// rcond = 1 / (||A||_1 ||A^{-1}||_1).  Per system the state word holds
// [0] the step whose result the next solve returns (kRc*), [1] the solve it
// needs (0: A^{-1}, 2: A^{-H}, -1: done), [2] iter, [3] j.
// ---------------------------------------------------------------------------

enum : int { kRcInit = 0, kRcSign1 = 1, kRcEj = 2, kRcSign2 = 3, kRcAlt = 4, kRcDone = 5 };

__global__ __launch_bounds__(256) void k_rcond_init(float *X, int64_t ldx, int n, int N,
                                                    int32_t *state, double *est) {
    const int b = blockIdx.x;
    float *xb = X + (size_t)b * ldx * 2;
    for (int i = threadIdx.x; i < N; i += 256)
        st(xb, i, c32{i < n ? 1.0f / (float)n : 0.0f, 0.0f});
    if (threadIdx.x == 0) {
        state[4 * b + 0] = kRcInit;
        state[4 * b + 1] = 0;
        state[4 * b + 2] = 0;
        state[4 * b + 3] = 0;
        est[b] = 0.0;
    }
}

// Y = X for the solve of this round (every system; the step ignores the
// systems that did not ask for it).
__global__ __launch_bounds__(256) void k_rcond_copy(const float *X, float *Y, int64_t ldx, int N) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) st(Y + (size_t)b * ldx * 2, i, ld(X + (size_t)b * ldx * 2, i));
}

// Consumes the solve result Y of the systems whose pending solve is
// `trans` and prepares their next input in X (clacn2 jumps 1-5).
__global__ __launch_bounds__(256) void k_rcond_step(const float *Y, float *X, int64_t ldx, int n,
                                                    int N, int trans, int32_t *state,
                                                    double *est) {
    const int b = blockIdx.x;
    int32_t *sw = state + 4 * b;
    if (sw[1] != trans) return;  // uniform over the block
    const float *yb = Y + (size_t)b * ldx * 2;
    float *xb = X + (size_t)b * ldx * 2;
    const int tid = threadIdx.x;
    __shared__ double sum_s[256];
    __shared__ float max_s[256];
    __shared__ int arg_s[256];
    __shared__ int next_s[2];  // next state, next solve
    // sum |y_i| (scsum1) and the first index of the largest |y_i| (icmax1)
    double sum = 0.0;
    float mx = -1.0f;
    int arg = 0;
    for (int i = tid; i < n; i += 256) {
        const c32 y = ld(yb, i);
        const float a = sqrtf(y.re * y.re + y.im * y.im);
        sum += (double)a;
        if (a > mx) {
            mx = a;
            arg = i;
        }
    }
    sum_s[tid] = sum;
    max_s[tid] = mx;
    arg_s[tid] = arg;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            sum_s[tid] += sum_s[tid + off];
            const float m2 = max_s[tid + off];
            const int a2 = arg_s[tid + off];
            if (m2 > max_s[tid] || (m2 == max_s[tid] && a2 < arg_s[tid])) {
                max_s[tid] = m2;
                arg_s[tid] = a2;
            }
        }
        __syncthreads();
    }
    const double l1 = sum_s[0];
    const int jmax = arg_s[0];
    // next input: 0 sign(y), 1 e_j, 2 alternating vector, -1 none
    int fill = -1, jsel = 0;
    if (tid == 0) {
        int stt = sw[0], op = -1;
        switch (stt) {
        case kRcInit:
            if (n == 1) {
                const c32 y = ld(yb, 0);
                est[b] = sqrt((double)y.re * y.re + (double)y.im * y.im);
                stt = kRcDone;
            } else {
                est[b] = l1;
                fill = 0;
                op = 2;
                stt = kRcSign1;
            }
            break;
        case kRcSign1:
            sw[3] = jmax;
            sw[2] = 2;
            fill = 1;
            op = 0;
            stt = kRcEj;
            break;
        case kRcEj: {
            const double old = est[b];
            est[b] = l1;
            if (l1 <= old) {
                fill = 2;
                op = 0;
                stt = kRcAlt;
            } else {
                fill = 0;
                op = 2;
                stt = kRcSign2;
            }
            break;
        }
        case kRcSign2: {
            const int jlast = sw[3];
            const c32 ya = ld(yb, jlast), yb2 = ld(yb, jmax);
            const float al = sqrtf(ya.re * ya.re + ya.im * ya.im);
            const float aj = sqrtf(yb2.re * yb2.re + yb2.im * yb2.im);
            sw[3] = jmax;
            if (al != aj && sw[2] < 5) {
                sw[2] += 1;
                fill = 1;
                op = 0;
                stt = kRcEj;
            } else {
                fill = 2;
                op = 0;
                stt = kRcAlt;
            }
            break;
        }
        case kRcAlt: {
            const double temp = 2.0 * (l1 / (3.0 * n));
            if (temp > est[b]) est[b] = temp;
            stt = kRcDone;
            break;
        }
        default:
            break;
        }
        sw[0] = stt;
        sw[1] = op;
        next_s[0] = fill;
        next_s[1] = sw[3];
    }
    __syncthreads();
    fill = next_s[0];
    jsel = next_s[1];
    if (fill < 0) return;
    const float safmin = 1.17549435e-38f;
    for (int i = tid; i < N; i += 256) {
        c32 v{0.0f, 0.0f};
        if (i < n) {
            if (fill == 0) {
                const c32 y = ld(yb, i);
                const float a = sqrtf(y.re * y.re + y.im * y.im);
                v = a > safmin ? c32{y.re / a, y.im / a} : c32{1.0f, 0.0f};
            } else if (fill == 1) {
                v.re = (i == jsel) ? 1.0f : 0.0f;
            } else {
                const float mag = 1.0f + (n > 1 ? (float)i / (float)(n - 1) : 0.0f);
                v.re = (i % 2 == 0) ? mag : -mag;
            }
        }
        st(xb, i, v);
    }
}

// ---------------------------------------------------------------------------
// FP64 products with the FP32 operator (true n x n block), batched over
// systems.  Workgroup = 64 rows of one system; its 4 waves take every 4th
// column of each 256-column chunk of x (staged through LDS), and their
// partial sums are added in wave order, so results do not depend on timing.
// ---------------------------------------------------------------------------

// (A x)_row for row = blockIdx.x * 64 + lane, valid in wave 0 of the block.
__device__ __forceinline__ void gemv_rows_f64(const float *Ab, int64_t lda, int n, int row,
                                              const double *xb, double &yr, double &yi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ double xsh[2][256];
    __shared__ double red[2][4][64];
    double ar = 0.0, ai = 0.0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        __syncthreads();
        const int c = c0 + threadIdx.x;
        xsh[0][threadIdx.x] = c < n ? xb[2 * c] : 0.0;
        xsh[1][threadIdx.x] = c < n ? xb[2 * c + 1] : 0.0;
        __syncthreads();
        if (row < n) {
            const int cend = min(256, n - c0);
#pragma unroll 8
            for (int cc = wave; cc < cend; cc += 4) {
                const c32 a = ld(Ab, (size_t)(c0 + cc) * lda + row);
                const double xr = xsh[0][cc], xi = xsh[1][cc];
                ar = fma((double)a.re, xr, ar);
                ar = fma(-(double)a.im, xi, ar);
                ai = fma((double)a.re, xi, ai);
                ai = fma((double)a.im, xr, ai);
            }
        }
    }
    red[0][wave][lane] = ar;
    red[1][wave][lane] = ai;
    __syncthreads();
    yr = ((red[0][0][lane] + red[0][1][lane]) + red[0][2][lane]) + red[0][3][lane];
    yi = ((red[1][0][lane] + red[1][1][lane]) + red[1][2][lane]) + red[1][3][lane];
}

// FP64 residual r = b - A x and the per-system maxima of |r| and |b|
// (norms[2 b], norms[2 b + 1], zero on entry).
__global__ __launch_bounds__(256) void k_residual_f64(const float *A, int64_t lda,
                                                      int64_t a_stride, int n, const float *bvec,
                                                      int64_t ldb, const double *x, int64_t ldx,
                                                      double *r, double *norms, int vpm) {
    const int b = blockIdx.y;
    const int row = blockIdx.x * 64 + (threadIdx.x & 63);
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    double ar, ai;
    gemv_rows_f64(Ab, lda, n, row, x + (size_t)b * ldx * 2, ar, ai);
    if (threadIdx.x >= 64) return;
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
    if (threadIdx.x == 0) {
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b),
                  static_cast<unsigned long long>(__double_as_longlong(rn)));
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b + 1),
                  static_cast<unsigned long long>(__double_as_longlong(bn)));
    }
}

// ---------------------------------------------------------------------------
// FP64 Krylov helpers of the GMRES refinement (batched over systems,
// vectors interleaved complex double with stride ld per system; the basis
// V holds vectors [b][i][ld]).
// ---------------------------------------------------------------------------

// y = A x, A the FP32 operator (true n x n block), x and y FP64.
__global__ __launch_bounds__(256) void k_gemv_f64(const float *A, int64_t lda, int64_t a_stride,
                                                  int n, const double *x, int64_t ldx, double *y,
                                                  int64_t ldy, int vpm) {
    const int b = blockIdx.y;
    const int row = blockIdx.x * 64 + (threadIdx.x & 63);
    const float *Ab = A + (size_t)(b / vpm) * a_stride * 2;
    double ar, ai;
    gemv_rows_f64(Ab, lda, n, row, x + (size_t)b * ldx * 2, ar, ai);
    if (threadIdx.x < 64 && row < n) {
        y[((size_t)b * ldy + row) * 2] = ar;
        y[((size_t)b * ldy + row) * 2 + 1] = ai;
    }
}

// h[b][i] = V_i^H w for i < m (grid.x = i, grid.y = b); h must be zero.
__global__ __launch_bounds__(256) void k_dots_f64(const double *V, int64_t ldv, int64_t vstride,
                                                  const double *w, int64_t ldw, int n, double *h,
                                                  int64_t ldh) {
    const int i = blockIdx.x, b = blockIdx.y;
    const double *v = V + ((size_t)b * vstride + (size_t)i * ldv) * 2;
    const double *wb = w + (size_t)b * ldw * 2;
    double re = 0.0, im = 0.0;
    for (int r = threadIdx.x; r < n; r += 256) {
        const double vr = v[2 * r], vi = v[2 * r + 1], wr = wb[2 * r], wi = wb[2 * r + 1];
        re = fma(vr, wr, fma(vi, wi, re));
        im = fma(vr, wi, fma(-vi, wr, im));
    }
    for (int o = 32; o >= 1; o >>= 1) {
        re += __shfl_xor(re, o);
        im += __shfl_xor(im, o);
    }
    __shared__ double red[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = re;
        red[1][wave] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        h[((size_t)b * ldh + i) * 2] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        h[((size_t)b * ldh + i) * 2 + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// w += sign * sum_{i < m} c[b][i] V_i  (sign = -1: Gram-Schmidt, +1: x += V y);
// systems with active[b] == 0 are left unchanged (active may be null).
__global__ __launch_bounds__(256) void k_combine_f64(const double *V, int64_t ldv, int64_t vstride,
                                                     int m, const double *c, int64_t ldc,
                                                     double sign, double *w, int64_t ldw, int n,
                                                     const int32_t *active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < m; ++i) {
        const double *v = V + ((size_t)b * vstride + (size_t)i * ldv) * 2;
        const double cr = c[((size_t)b * ldc + i) * 2], ci = c[((size_t)b * ldc + i) * 2 + 1];
        const double vr = v[2 * r], vi = v[2 * r + 1];
        re = fma(cr, vr, fma(-ci, vi, re));
        im = fma(cr, vi, fma(ci, vr, im));
    }
    double *wb = w + (size_t)b * ldw * 2;
    wb[2 * r] += sign * re;
    wb[2 * r + 1] += sign * im;
}

// out[b] = ||w_b||_2^2 (out must be zero).
__global__ __launch_bounds__(256) void k_norm2sq_f64(const double *w, int64_t ldw, int n,
                                                     double *out) {
    const int b = blockIdx.y;
    const double *wb = w + (size_t)b * ldw * 2;
    double s = 0.0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256)
        s = fma(wb[2 * r], wb[2 * r], fma(wb[2 * r + 1], wb[2 * r + 1], s));
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(out + b, s);
}

// dst_b = src_b * f[b] (real factor per system); dst may alias src.
__global__ void k_scale_f64(const double *src, int64_t lds, double *dst, int64_t ldd, int n,
                            const double *f) {
    const int b = blockIdx.y;
    const double s = f[b];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < 2 * n; i += gridDim.x * 256)
        dst[(size_t)b * ldd * 2 + i] = src[(size_t)b * lds * 2 + i] * s;
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

void launch_pad_identity_var(float *A, int64_t lda, int64_t a_stride, const int32_t *n_sys, int N,
                             int B, hipStream_t stream) {
    hipLaunchKernelGGL(k_pad_identity_var, dim3(N, B), dim3(256), 0, stream, A, lda, a_stride, n_sys,
                       N);
}

void launch_norm1(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *norms,
                  double *amax, hipStream_t stream) {
    launch_fill_zero_f64(norms, (size_t)B, stream);
    launch_fill_zero_f64(amax, (size_t)B, stream);
    hipLaunchKernelGGL(k_norm1, dim3((n + 15) / 16, B), dim3(256), 0, stream, A, lda, a_stride, n,
                       norms, amax);
}

void launch_pow2_scale(const double *amax, float *scale, int B, hipStream_t stream) {
    hipLaunchKernelGGL(k_pow2_scale, dim3((B + 255) / 256), dim3(256), 0, stream, amax, scale, B);
}

// Two-level blocked LU: outer blocks of kLuOuter columns are factored with
// kNB-wide panels, their exchanges, unit-lower solves and rank-kNB updates
// (FP32) inside the block; the rows of the block right of it then get all
// of the block's exchanges, the block-triangular solve (kNB rows at a time
// with FP32 rank-kNB updates of the block's remaining rows), and one
// rank-kLuOuter update of the trailing matrix in the requested precision.
void lu_factor(const LuWorkspace &w, int precision, const float *scale, hipStream_t stream) {
    static_assert(kLuPad % kNB == 0 && kLuOuter % kNB == 0, "whole panels");
    static_assert(kLuOuter % kTrK == 0 && kLuOuter % kTr32K == 0, "whole trailing-update stages");
    if (w.N > kLuMaxOrder) throw std::runtime_error("LU order above the supported maximum");
    if (precision != 0 && w.N > kLuOuter && !w.pack)
        throw std::runtime_error("16-bit LU needs the packed-operand workspace");
    // int32 zero has the bit pattern of float zero
    launch_fill_zero(reinterpret_cast<float *>(w.info), (size_t)w.B, stream);
    const int N = w.N, B = w.B;
    const int64_t gls = (int64_t)(N / kNB) * kLuGather;
    auto gl_of = [&](int k) { return w.gl + (size_t)(k / kNB) * kLuGather; };
    // exchanges / solves of step k (n_steps exchanges for mode 0) on up to
    // two column ranges in one launch
    auto swap_trsm = [&](int k, int n_steps, int a0, int a1, int mode_a, int b0, int b1,
                         int mode_b) {
        a1 = std::max(a0, a1);
        b1 = std::max(b0, b1);
        const int blocks = (a1 - a0 + 63) / 64 + (b1 - b0 + 63) / 64;
        if (blocks == 0) return;
        hipLaunchKernelGGL(k_lu_swap_trsm, dim3(blocks, B), dim3(256), 0, stream, w.A, w.lda,
                           w.a_stride, k, n_steps, a0, a1, mode_a, b0, b1, mode_b, gl_of(k), gls);
    };
    auto gemm = [&](int r0, int r1, int c0, int c1, int k0) {
        if (r1 <= r0 || c1 <= c0) return;
        const dim3 grid((r1 - r0 + 63) / 64, (c1 - c0 + 63) / 64, B);
        hipLaunchKernelGGL(k_lu_gemm, grid, dim3(256), 0, stream, w.A, w.lda, w.a_stride, r0, r1,
                           c0, c1, k0);
    };
    for (int K0 = 0; K0 < N; K0 += kLuOuter) {
        const int K1 = std::min(N, K0 + kLuOuter);
        for (int k = K0; k < K1; k += kNB) {
            if (N - k <= 1024) {
                const int threads = (N - k + 63) / 64 * 64;
                hipLaunchKernelGGL(k_lu_panel_reg, dim3(B), dim3(threads), 0, stream, w.A, w.lda,
                                   w.a_stride, N, k, w.ipiv, w.info, gl_of(k), gls);
            } else {
                const int G = (N - k + 1023) / 1024;
                hipLaunchKernelGGL(k_lu_panel_nominate, dim3(G, B), dim3(1024), 0, stream, w.A,
                                   w.lda, w.a_stride, N, k, w.scratch_i);
                // one thread per candidate row (whole waves)
                const int tthreads = std::min(1024, (G * kNB + 63) / 64 * 64);
                hipLaunchKernelGGL(k_lu_panel_tournament, dim3(B), dim3(tthreads), 0, stream, w.A,
                                   w.lda, w.a_stride, N, k, G, w.scratch_i, w.ipiv, w.info,
                                   gl_of(k), gls, w.scratch_f, w.scratch_f + (size_t)B * kNB * kNB * 2);
                hipLaunchKernelGGL(k_lu_panel_apply, dim3((N - k + 255) / 256, B), dim3(256), 0,
                                   stream, w.A, w.lda, w.a_stride, N, k, gl_of(k), gls,
                                   w.scratch_f, w.scratch_f + (size_t)B * kNB * kNB * 2);
            }
            swap_trsm(k, 1, K0, k, 0, k + kNB, K1, 1);
            gemm(k + kNB, N, k + kNB, K1, k);
        }
        if (K1 < N) {
            const int br_tiles = (N - K1 + kBrCols - 1) / kBrCols;
            if (N <= kBrMaxOrder && (int64_t)br_tiles * B <= kBrMaxGroups) {
                static const bool lds_attr = [] {
                    // dynamic LDS beside the static block row (above 64 KiB in all)
                    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lu_block_row),
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)(sizeof(int) * kBrMaxOrder));
                    return true;
                }();
                (void)lds_attr;
                const dim3 grid(br_tiles, B);
                hipLaunchKernelGGL(k_lu_block_row, grid, dim3(256), sizeof(int) * (size_t)(N - K0),
                                   stream, w.A, w.lda, w.a_stride, N, K0, K1, w.gl, gls);
            } else {
                swap_trsm(K0, (K1 - K0) / kNB, K1, N, 0, 0, 0, 0);
                for (int k = K0; k < K1; k += kNB) {
                    swap_trsm(k, 1, K1, N, 2, 0, 0, 0);
                    gemm(k + kNB, K1, K1, N, k);
                }
            }
            if (precision != 0) {
                // 16-bit operands packed once, then the 128-tile MFMA update
                const int M = N - K1;
                const int64_t plane = (int64_t)N * kLuOuter, pstride = 4 * plane;
                if (precision == 1) {
                    hipLaunchKernelGGL(k_lu_pack_l<1>, dim3((M + 63) / 64, kLuOuter / 32, B),
                                       dim3(256), 0, stream, w.A, w.lda, w.a_stride, K0, K1, M,
                                       w.pack, pstride, plane);
                    hipLaunchKernelGGL(k_lu_pack_u<1>, dim3((M + 7) / 8, B), dim3(256), 0, stream,
                                       w.A, w.lda, w.a_stride, K0, K1, M, w.pack, pstride, plane,
                                       scale);
                } else {
                    hipLaunchKernelGGL(k_lu_pack_l<2>, dim3((M + 63) / 64, kLuOuter / 32, B),
                                       dim3(256), 0, stream, w.A, w.lda, w.a_stride, K0, K1, M,
                                       w.pack, pstride, plane);
                    hipLaunchKernelGGL(k_lu_pack_u<2>, dim3((M + 7) / 8, B), dim3(256), 0, stream,
                                       w.A, w.lda, w.a_stride, K0, K1, M, w.pack, pstride, plane,
                                       scale);
                }
                const dim3 grid((M + kTrTile - 1) / kTrTile, (M + kTrTile - 1) / kTrTile, B);
                if (precision == 1)
                    hipLaunchKernelGGL(k_lu_gemm_trp<1>, grid, dim3(256), 0, stream, w.A, w.lda,
                                       w.a_stride, K1, M, w.pack, pstride, plane, scale);
                else
                    hipLaunchKernelGGL(k_lu_gemm_trp<2>, grid, dim3(256), 0, stream, w.A, w.lda,
                                       w.a_stride, K1, M, w.pack, pstride, plane, scale);
            } else {
                // the rank-kLuOuter trailing update in FP32 on the 128-tile kernel
                const dim3 grid((N - K1 + kTrTile - 1) / kTrTile, (N - K1 + kTrTile - 1) / kTrTile, B);
                hipLaunchKernelGGL(k_lu_gemm_tr32, grid, dim3(256), 0, stream, w.A, w.lda,
                                   w.a_stride, K1, N, K1, N, K0, K1);
            }
        }
    }
    if (w.perm && w.pi && w.dinv && w.lf && N <= kTrsvMaxOrder) {
        const size_t shm = sizeof(int32_t) * 2 * (size_t)N;
        hipLaunchKernelGGL(k_lu_solve_tables, dim3(B), dim3(1024), shm, stream, w.gl, gls, N,
                           w.perm, w.pi);
        hipLaunchKernelGGL(k_lu_final_l, dim3(N, B), dim3(256), 0, stream, w.A, w.lda, w.a_stride,
                           N, w.pi, w.lf);
        hipLaunchKernelGGL(k_lu_diag_inverse, dim3(N / kTrsvRows, B, 2), dim3(64), 0, stream, w.A,
                           w.lda, w.a_stride, N, w.dinv);
    }
}

void launch_trsv_init(const TrsvSync &sync, size_t n_words, hipStream_t stream) {
    const size_t n = std::max<size_t>(4, n_words);
    hipLaunchKernelGGL(k_trsv_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       sync.ctx, sync.y, sync.z, n_words);
}

void lu_solve(const LuWorkspace &w, float *x, int64_t ldx, int trans, hipStream_t stream,
              int vpm, const TrsvSync *sync) {
    const int N = w.N, V = w.B * vpm;
    if (trans != 0 && trans != 2) throw std::runtime_error("lu_solve: trans must be 0 or 2");
    const int64_t gls = (int64_t)(N / kNB) * kLuGather;
    if (sync && w.perm && w.pi && w.dinv && w.lf && N <= kTrsvMaxOrder) {
        const dim3 grid((unsigned)(V * (N / kTrsvRows)));
        if (trans == 0) {
            hipLaunchKernelGGL(k_trsv<0>, grid, dim3(256), 0, stream, w.A, w.lf, w.lda, w.a_stride,
                               N, w.perm, w.dinv, x, ldx, vpm, *sync);
            hipLaunchKernelGGL(k_trsv<1>, grid, dim3(256), 0, stream, w.A, w.lf, w.lda, w.a_stride,
                               N, w.perm, w.dinv, x, ldx, vpm, *sync);
        } else {
            hipLaunchKernelGGL(k_trsv<2>, grid, dim3(256), 0, stream, w.A, w.lf, w.lda, w.a_stride,
                               N, w.perm, w.dinv, x, ldx, vpm, *sync);
            hipLaunchKernelGGL(k_trsv<3>, grid, dim3(256), 0, stream, w.A, w.lf, w.lda, w.a_stride,
                               N, w.perm, w.dinv, x, ldx, vpm, *sync);
        }
    } else if (N <= kLuSolveOneBlock) {
        const size_t shm = sizeof(float) * 2 * (size_t)N;
        hipLaunchKernelGGL(k_lu_solve, dim3(V), dim3(1024), shm, stream, w.A, w.lda, w.a_stride,
                           N, w.ipiv, w.gl, gls, x, ldx, trans, vpm);
    } else {
        auto blk = [&](int K0, int mode) {
            hipLaunchKernelGGL(k_solve_block, dim3(V), dim3(256), 0, stream, w.A, w.lda,
                               w.a_stride, N, w.gl, gls, x, ldx, K0, mode, vpm);
        };
        auto upd_n = [&](int K0, int K1, int r0, int r1) {
            if (r1 > r0)
                hipLaunchKernelGGL(k_solve_upd_n, dim3((r1 - r0 + 255) / 256, V), dim3(256), 0,
                                   stream, w.A, w.lda, w.a_stride, x, ldx, K0, K1, r0, r1, vpm);
        };
        auto upd_h = [&](int K0, int K1, int r0, int r1) {
            if (r1 > r0)
                hipLaunchKernelGGL(k_solve_upd_h, dim3((K1 - K0 + 3) / 4, V), dim3(256), 0,
                                   stream, w.A, w.lda, w.a_stride, x, ldx, K0, K1, r0, r1, vpm);
        };
        const int last = (N - 1) / kLuOuter * kLuOuter;
        if (trans == 0) {
            for (int K0 = 0; K0 < N; K0 += kLuOuter) {
                const int K1 = std::min(N, K0 + kLuOuter);
                blk(K0, 0);
                upd_n(K0, K1, K1, N);
            }
            for (int K0 = last; K0 >= 0; K0 -= kLuOuter) {
                const int K1 = std::min(N, K0 + kLuOuter);
                blk(K0, 1);
                upd_n(K0, K1, 0, K0);
            }
        } else {
            for (int K0 = 0; K0 < N; K0 += kLuOuter) {
                const int K1 = std::min(N, K0 + kLuOuter);
                upd_h(K0, K1, 0, K0);
                blk(K0, 2);
            }
            for (int K0 = last; K0 >= 0; K0 -= kLuOuter) {
                const int K1 = std::min(N, K0 + kLuOuter);
                upd_h(K0, K1, K1, N);
                blk(K0, 3);
            }
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("LU solve launch: ") + hipGetErrorString(e));
}

void launch_rcond_init(float *X, int64_t ldx, int n, int N, int B, int32_t *state, double *est,
                       hipStream_t stream) {
    hipLaunchKernelGGL(k_rcond_init, dim3(B), dim3(256), 0, stream, X, ldx, n, N, state, est);
}

void launch_rcond_copy(const float *X, float *Y, int64_t ld, int N, int B, hipStream_t stream) {
    hipLaunchKernelGGL(k_rcond_copy, dim3((N + 255) / 256, B), dim3(256), 0, stream, X, Y, ld, N);
}

void launch_rcond_step(const float *Y, float *X, int64_t ldx, int n, int N, int B, int trans,
                       int32_t *state, double *est, hipStream_t stream) {
    hipLaunchKernelGGL(k_rcond_step, dim3(B), dim3(256), 0, stream, Y, X, ldx, n, N, trans, state,
                       est);
}

void launch_residual_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                         const float *b, int64_t ldb, const double *x, int64_t ldx, double *r,
                         double *norms, hipStream_t stream, int vpm) {
    launch_fill_zero_f64(norms, (size_t)2 * B, stream);
    hipLaunchKernelGGL(k_residual_f64, dim3((n + 63) / 64, B), dim3(256), 0, stream, A, lda,
                       a_stride, n, b, ldb, x, ldx, r, norms, vpm);
}

void launch_gemv_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B, const double *x,
                     int64_t ldx, double *y, int64_t ldy, hipStream_t stream, int vpm) {
    hipLaunchKernelGGL(k_gemv_f64, dim3((n + 63) / 64, B), dim3(256), 0, stream, A, lda, a_stride,
                       n, x, ldx, y, ldy, vpm);
}

void launch_dots_f64(const double *V, int64_t ldv, int64_t vstride, int m, const double *w,
                     int64_t ldw, int n, int B, double *h, int64_t ldh, hipStream_t stream) {
    hipLaunchKernelGGL(k_dots_f64, dim3(m, B), dim3(256), 0, stream, V, ldv, vstride, w, ldw, n, h,
                       ldh);
}

void launch_combine_f64(const double *V, int64_t ldv, int64_t vstride, int m, const double *c,
                        int64_t ldc, double sign, double *w, int64_t ldw, int n, int B,
                        const int32_t *active, hipStream_t stream) {
    hipLaunchKernelGGL(k_combine_f64, dim3((n + 255) / 256, B), dim3(256), 0, stream, V, ldv,
                       vstride, m, c, ldc, sign, w, ldw, n, active);
}

void launch_norm2sq_f64(const double *w, int64_t ldw, int n, int B, double *out,
                        hipStream_t stream) {
    launch_fill_zero_f64(out, (size_t)B, stream);
    hipLaunchKernelGGL(k_norm2sq_f64, dim3(std::min(64, (n + 255) / 256), B), dim3(256), 0, stream,
                       w, ldw, n, out);
}

void launch_scale_f64(const double *src, int64_t lds, double *dst, int64_t ldd, int n, int B,
                      const double *f, hipStream_t stream) {
    hipLaunchKernelGGL(k_scale_f64, dim3((2 * n + 255) / 256, B), dim3(256), 0, stream, src, lds,
                       dst, ldd, n, f);
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
