// gfx950 kernels of the CHIEF interior-point rows.
//
// For interior point i and P1 column j (assembleChiefRows,
// main.swift:1930-2036):
//   C[i][j] = sum_t sum_images int_t dG/dn_y (x_i, y) phi_j(y) dy
//             - (i k beta_t / 3) int_t G(x_i, y) dy         (Robin fold)
//   d_s[i]  = sum_t int_t G(x_i, y) dy  q_s[t]
// with the real wavenumber in G and dG/dn (the reference evaluates these
// rows with the field kernels) and i k = -k_im + i k_re in the fold.
// Lanes own triangles (device order); one block row per interior point and
// frequency; C and d are accumulated with float atomics.
#include <hip/hip_runtime.h>

#include "hbem_internal.h"

namespace hbem {
namespace {

constexpr float kInv4Pi = 0.07957747154594767f;
constexpr float kInv2Pi = 0.15915494309189535f;

__device__ __forceinline__ float msign(int mask, int bit) { return (mask & bit) ? -1.0f : 1.0f; }

__global__ __launch_bounds__(256) void k_chief_rows(DeviceGeometry g, int m,
                                                    const float *__restrict__ pts,
                                                    const float *__restrict__ k_real,
                                                    const float *__restrict__ k_imag, int n_src,
                                                    const float *__restrict__ q,
                                                    const float *__restrict__ beta,
                                                    float *__restrict__ C, int n,
                                                    float *__restrict__ d) {
    const int T = g.n_triangles;
    const int td = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    const int b = blockIdx.z;
    if (td >= T) return;
    const float *tf = g.tri_f_soa;
    const int32_t *ti = g.tri_i_soa;
    const float ox = pts[i], oy = pts[m + i], oz = pts[2 * m + i];
    const float k = k_real[b], krev = k * kInv2Pi;
    const float jac = tf[(size_t)22 * T + td];
    const float nx = tf[(size_t)18 * T + td], ny = tf[(size_t)19 * T + td],
                nz = tf[(size_t)20 * T + td];
    float dre[3] = {0, 0, 0}, dim[3] = {0, 0, 0}, gre = 0.0f, gim = 0.0f;
    for (int mi = 0; mi < g.n_masks; ++mi) {
        const int mask = g.masks[mi];
        const float sx = msign(mask, 1), sy = msign(mask, 2), sz = msign(mask, 4);
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const float w = (float)kRule6W[a] * jac;
            const float dx = sx * tf[(size_t)(3 * a) * T + td] - ox;
            const float dy = sy * tf[(size_t)(3 * a + 1) * T + td] - oy;
            const float dz = sz * tf[(size_t)(3 * a + 2) * T + td] - oz;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            float t = r * krev;
            t = t - __builtin_rintf(t);
            const float s = __builtin_amdgcn_sinf(t), c = __builtin_amdgcn_cosf(t);
            const float sc = kInv4Pi * rinv;
            const float g_re = c * sc, g_im = s * sc;
            const float proj = fmaf(dx, sx * nx, fmaf(dy, sy * ny, dz * sz * nz)) * rinv;
            // dG/dn = G (i k - 1/r) proj
            const float d_re = (g_re * -rinv - g_im * k) * proj;
            const float d_im = (g_re * k + g_im * -rinv) * proj;
            const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
            dre[0] = fmaf(d_re, b0 * w, dre[0]);
            dim[0] = fmaf(d_im, b0 * w, dim[0]);
            dre[1] = fmaf(d_re, xi * w, dre[1]);
            dim[1] = fmaf(d_im, xi * w, dim[1]);
            dre[2] = fmaf(d_re, eta * w, dre[2]);
            dim[2] = fmaf(d_im, eta * w, dim[2]);
            gre = fmaf(g_re, w, gre);
            gim = fmaf(g_im, w, gim);
        }
    }
    const int orig = ti[(size_t)19 * T + td];
    float fr = 0.0f, fi = 0.0f;
    if (beta) {
        const float br = beta[((size_t)b * T + orig) * 2], bi = beta[((size_t)b * T + orig) * 2 + 1];
        const float ikr = k_imag ? -k_imag[b] : 0.0f;
        const float cr = ikr * br - k * bi, ci = ikr * bi + k * br;  // i k beta
        fr = (gre * cr - gim * ci) * (1.0f / 3.0f);
        fi = (gre * ci + gim * cr) * (1.0f / 3.0f);
    }
    float *Cb = C + ((size_t)b * m + i) * n * 2;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const int col = ti[(size_t)l * T + td];
        atomicAdd(Cb + 2 * col, dre[l] - fr);
        atomicAdd(Cb + 2 * col + 1, dim[l] - fi);
    }
    for (int sv = 0; sv < n_src; ++sv) {
        const size_t qi = (((size_t)b * n_src + sv) * T + orig) * 2;
        const float qr = q[qi], qim = q[qi + 1];
        if (qr == 0.0f && qim == 0.0f) continue;
        float *db = d + (((size_t)b * n_src + sv) * m + i) * 2;
        atomicAdd(db, gre * qr - gim * qim);
        atomicAdd(db + 1, gre * qim + gim * qr);
    }
}

// ||A||_inf = max_i sum_j |a_ij| per matrix (column-major interleaved
// complex, FP64 sums); one thread per row, rows of a block contiguous.
__global__ __launch_bounds__(256) void k_row_abs_sum(const float *__restrict__ A, int64_t lda,
                                                     int64_t a_stride, int n,
                                                     double *__restrict__ out) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    double sum = 0.0;
    if (row < n) {
        const float *Ab = A + (size_t)b * a_stride * 2;
        for (int c = 0; c < n; ++c) {
            const double re = Ab[((size_t)c * lda + row) * 2], im = Ab[((size_t)c * lda + row) * 2 + 1];
            sum += sqrt(re * re + im * im);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) sum = fmax(sum, __shfl_xor(sum, o));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(out + b),
                  static_cast<unsigned long long>(__double_as_longlong(sum)));
}

}  // namespace

void launch_chief_rows(const DeviceGeometry &g, int n_batch, int m, const float *pts,
                       const float *k_real, const float *k_imag, int n_src, const float *q,
                       const float *beta, float *C, float *d, hipStream_t stream) {
    if (m <= 0 || n_batch <= 0) return;
    dim3 grid((g.n_triangles + 255) / 256, m, n_batch);
    hipLaunchKernelGGL(k_chief_rows, grid, dim3(256), 0, stream, g, m, pts, k_real, k_imag, n_src,
                       q, beta, C, g.n_dofs, d);
}

void launch_row_abs_sum(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *out,
                        hipStream_t stream) {
    launch_fill_zero_f64(out, (size_t)B, stream);
    hipLaunchKernelGGL(k_row_abs_sum, dim3((n + 255) / 256, B), dim3(256), 0, stream, A, lda,
                       a_stride, n, out);
}

}  // namespace hbem
