// gfx950 kernels of the adjoint (shape derivative) of the field pipeline.
//
// For a real loss L of the exterior field f = E p + F qt (qt = q + i k beta
// mean(p), the Neumann data the field sees) with A p = b = V q, and the
// complex cotangent g = dL/dRe f + i dL/dIm f:
//   lambda = A^{-H} (E^H g + R^H F^H g)      (R: p -> i k beta mean(p))
//   dL/dv  = d/dv Re(g^H (E p + F qt))  +  d/dv Re(lambda^H (V q - A p))
// with lambda, p and qt held fixed in the derivatives (the second term is
// the usual adjoint identity; the Robin fold of A turns V q - A p into the
// same V qt - K p form per pair).  Every pair block of the forward assembly
// (regular 6x6, Sauter-Schwab classes, mass term) and of the field is
// differentiated through its quadrature points, its Jacobian |N| and its
// normal N = +-(v1 - v0) x (v2 - v0), recomputing the points rather than
// storing them.  dL/dq = F^H g + V^H lambda comes from the same sums.
//
// Per point pair with r = y - x, G = e^{ikr}/(4 pi r) (complex k), u = i k:
//   grad_r G           = H r,              H  = G (u - 1/r) / r
//   grad_r H (r . N)   = (G Q / r^2)(r . N) r + H N,
//                                           Q  = u^2 - 3 u / r + 3 / r^2
// Point-position gradients are accumulated per vertex, normal gradients per
// triangle (k_adjoint_normals turns them into vertex gradients).
#include <hip/hip_runtime.h>

#include "hbem_internal.h"

namespace hbem {
namespace {

constexpr float kInv4Pi = 0.07957747154594767f;
constexpr float kInv2Pi = 0.15915494309189535f;
constexpr int kSgRun = 32;  // FP32 run length of the singular sums (rule counts are multiples)

struct AdjK {
    float gre, gim;    // G
    float hre, him;    // H
    float qre, qim;    // G Q / r^2
};

template <bool kComplexK>
__device__ __forceinline__ AdjK adj_kernel(float dx, float dy, float dz, float k, float krev,
                                           float kim) {
    const float r2raw = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool ok = r2raw > 1.0e-14f;
    const float r2 = ok ? r2raw : 1.0e-14f;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float t = r * krev;
    t = t - __builtin_rintf(t);
    const float s = __builtin_amdgcn_sinf(t);
    const float c = __builtin_amdgcn_cosf(t);
    float sc = kInv4Pi * rinv;
    if (kComplexK) sc *= __expf(-kim * r);
    sc = ok ? sc : 0.0f;
    AdjK o;
    o.gre = c * sc;
    o.gim = s * sc;
    const float fre = kComplexK ? (-kim - rinv) : -rinv;
    o.hre = (o.gre * fre - o.gim * k) * rinv;
    o.him = (o.gre * k + o.gim * fre) * rinv;
    // Q = u^2 - 3 u / r + 3 / r^2 with u = i k - k_im
    const float r2inv = rinv * rinv;
    const float qr = kComplexK ? fmaf(kim, kim, -k * k) + 3.0f * kim * rinv + 3.0f * r2inv
                               : 3.0f * r2inv - k * k;
    const float qi = kComplexK ? -2.0f * k * kim - 3.0f * k * rinv : -3.0f * k * rinv;
    o.qre = (o.gre * qr - o.gim * qi) * r2inv;
    o.qim = (o.gre * qi + o.gim * qr) * r2inv;
    return o;
}

__device__ __forceinline__ float msign(int mask, int bit) { return (mask & bit) ? -1.0f : 1.0f; }

constexpr double kInv2PiD = 0.15915494309189533577;

// adj_kernel for an observation point far from the surface: d given in
// FP64, distance and phase in FP64, the rest in FP32.
__device__ __forceinline__ AdjK adj_kernel_far(double dx, double dy, double dz, float k,
                                               double krev) {
    const double r2d = fma(dx, dx, fma(dy, dy, dz * dz));
    const bool ok = r2d > 1.0e-14;
    const float r2 = ok ? (float)r2d : 1.0f;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r32 = r2 * rinv;
    const double rd = (double)r32 + fma(-(double)r32, (double)r32, r2d) * (double)(0.5f * rinv);
    double td = rd * krev;
    td -= rint(td);
    const float t = (float)td;
    const float s = __builtin_amdgcn_sinf(t);
    const float c = __builtin_amdgcn_cosf(t);
    const float sc = ok ? kInv4Pi * rinv : 0.0f;
    AdjK o;
    o.gre = c * sc;
    o.gim = s * sc;
    o.hre = (-o.gre * rinv - o.gim * k) * rinv;
    o.him = (o.gre * k - o.gim * rinv) * rinv;
    const float r2inv = rinv * rinv;
    const float qr = 3.0f * r2inv - k * k, qi = -3.0f * k * rinv;
    o.qre = (o.gre * qr - o.gim * qi) * r2inv;
    o.qim = (o.gre * qi + o.gim * qr) * r2inv;
    return o;
}

// adj_kernel with the distance and the phase k r formed in FP64 from the
// components of d (which the regular adjoint forms without cancellation,
// from vertex-anchored offsets): the phase error no longer grows with k r.
template <bool kComplexK>
__device__ __forceinline__ AdjK adj_kernel_p(double dx, double dy, double dz, float k, double krev,
                                             float kim) {
    const double r2d = fma(dx, dx, fma(dy, dy, dz * dz));
    const bool ok = r2d > 1.0e-14;
    const float r2 = ok ? (float)r2d : 1.0e-14f;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r32 = r2 * rinv;
    const double rd = (double)r32 + fma(-(double)r32, (double)r32, r2d) * (double)(0.5f * rinv);
    double td = rd * krev;
    td -= rint(td);
    const float t = (float)td;
    const float r = (float)rd;
    const float s = __builtin_amdgcn_sinf(t);
    const float c = __builtin_amdgcn_cosf(t);
    float sc = kInv4Pi * rinv;
    if (kComplexK) sc *= __expf(-kim * r);
    sc = ok ? sc : 0.0f;
    AdjK o;
    o.gre = c * sc;
    o.gim = s * sc;
    const float fre = kComplexK ? (-kim - rinv) : -rinv;
    o.hre = (o.gre * fre - o.gim * k) * rinv;
    o.him = (o.gre * k + o.gim * fre) * rinv;
    const float r2inv = rinv * rinv;
    const float qr = kComplexK ? fmaf(kim, kim, -k * k) + 3.0f * kim * rinv + 3.0f * r2inv
                               : 3.0f * r2inv - k * k;
    const float qi = kComplexK ? -2.0f * k * kim - 3.0f * k * rinv : -3.0f * k * rinv;
    o.qre = (o.gre * qr - o.gim * qi) * r2inv;
    o.qim = (o.gre * qi + o.gim * qr) * r2inv;
    return o;
}

__device__ __forceinline__ void atomic_add3(double *g, int stride, int idx, double x, double y,
                                            double z) {
    atomicAdd(g + idx, x);
    atomicAdd(g + stride + idx, y);
    atomicAdd(g + 2 * stride + idx, z);
}

// ---------------------------------------------------------------------------
// Field adjoint.  Lanes own triangles (device order), the wave walks a
// chunk of observation points (uniform loads).  For every frequency (grid.z)
// and triangle t, over points a and images m with r = y_a - S_m x_o (trial
// frame, N = N_t):
//   c_j  += conj(g_o) w_a phi_j(a) H (r . N) P_a-free part   (-> E^H g)
//   ft_t += -conj(g_o) w_a |N_t| G                          (-> F^H g)
// and the gradients of Re(conj(g_o) w_a [H (r . N) P_a - |N_t| G qt_t]).
// ---------------------------------------------------------------------------
constexpr int kFieldAdjChunk = 64;

__global__ __launch_bounds__(256) void k_field_adjoint(
    DeviceGeometry g, const float *__restrict__ field_k, const float *__restrict__ p,
    int64_t p_stride, const float *__restrict__ qt, int n_obs, const float *__restrict__ obs,
    const float *__restrict__ grad_field, double *__restrict__ c_out, int64_t c_stride,
    double *__restrict__ ft_out, double *__restrict__ gV, double *__restrict__ gN) {
    const int T = g.n_triangles, nv = g.n_vertices;
    const int td = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.z;
    const bool live = td < T;
    const int tdd = live ? td : T - 1;
    const float *tf = g.tri_f_soa;
    const int32_t *ti = g.tri_i_soa;
    // quadrature points in FP64 from the (exact float) vertices
    double y[18];
    {
        double v0[3], e1[3], e2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v0[c] = tf[(size_t)(23 + c) * T + tdd];
            e1[c] = (double)tf[(size_t)(26 + c) * T + tdd] - v0[c];
            e2[c] = (double)tf[(size_t)(29 + c) * T + tdd] - v0[c];
        }
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) y[3 * a + c] = fma(kRule6X[a], e1[c], fma(kRule6Y[a], e2[c], v0[c]));
    }
    const float jac = tf[(size_t)22 * T + tdd];
    const float nx = tf[(size_t)18 * T + tdd] * jac, ny = tf[(size_t)19 * T + tdd] * jac,
                nz = tf[(size_t)20 * T + tdd] * jac;
    const float nhx = tf[(size_t)18 * T + tdd], nhy = tf[(size_t)19 * T + tdd],
                nhz = tf[(size_t)20 * T + tdd];
    int dof[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) dof[l] = ti[(size_t)l * T + tdd];
    const int orig = ti[(size_t)19 * T + tdd];
    const float *pb = p + (size_t)b * p_stride * 2;
    float pr[3], pim[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        pr[l] = pb[2 * dof[l]];
        pim[l] = pb[2 * dof[l] + 1];
    }
    const float qr = qt[((size_t)b * T + orig) * 2], qi = qt[((size_t)b * T + orig) * 2 + 1];
    float Pr[6], Pi[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
        Pr[a] = b0 * pr[0] + xi * pr[1] + eta * pr[2];
        Pi[a] = b0 * pim[0] + xi * pim[1] + eta * pim[2];
    }
    const float k = field_k[b];
    const double krev = (double)k * kInv2PiD;
    const float *gb = grad_field + (size_t)b * n_obs * 2;

    double cj[3][2] = {{0, 0}, {0, 0}, {0, 0}}, ftd[2] = {0, 0};
    double gy[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gn[3] = {0, 0, 0};
    const int o_begin = blockIdx.y * kFieldAdjChunk;
    const int o_end = min(n_obs, o_begin + kFieldAdjChunk);
    for (int o = o_begin; o < o_end; ++o) {
        const float ox = obs[o], oy = obs[n_obs + o], oz = obs[2 * n_obs + o];
        // gamma = conj(g_o)
        const float gr = gb[2 * o], gi = -gb[2 * o + 1];
        float c_re[3] = {0, 0, 0}, c_im[3] = {0, 0, 0}, f_re = 0.0f, f_im = 0.0f;
        float gya[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gna[3] = {0, 0, 0};
        for (int m = 0; m < g.n_masks; ++m) {
            const int mask = g.masks[m];
            const float sx = msign(mask, 1), sy = msign(mask, 2), sz = msign(mask, 4);
            const double mx = sx * ox, my = sy * oy, mz = sz * oz;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const float w = (float)kRule6W[a];
                const double rxd = y[3 * a] - mx, ryd = y[3 * a + 1] - my, rzd = y[3 * a + 2] - mz;
                const AdjK K = adj_kernel_far(rxd, ryd, rzd, k, krev);
                const float rx = (float)rxd, ry = (float)ryd, rz = (float)rzd;
                const float rn = fmaf(rx, nx, fmaf(ry, ny, rz * nz));
                // psi = H (r . N);  gamma psi
                const float psr = K.hre * rn, psi = K.him * rn;
                const float gpr = (gr * psr - gi * psi) * w, gpi = (gr * psi + gi * psr) * w;
                const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
                c_re[0] = fmaf(gpr, b0, c_re[0]);
                c_im[0] = fmaf(gpi, b0, c_im[0]);
                c_re[1] = fmaf(gpr, xi, c_re[1]);
                c_im[1] = fmaf(gpi, xi, c_im[1]);
                c_re[2] = fmaf(gpr, eta, c_re[2]);
                c_im[2] = fmaf(gpi, eta, c_im[2]);
                // gamma G |N| w (F part carries a minus sign)
                const float ggr = (gr * K.gre - gi * K.gim) * w, ggi = (gr * K.gim + gi * K.gre) * w;
                f_re = fmaf(-jac, ggr, f_re);
                f_im = fmaf(-jac, ggi, f_im);
                // beta = gamma P_a w, alpha = gamma qt w
                const float br = (gr * Pr[a] - gi * Pi[a]) * w, bi = (gr * Pi[a] + gi * Pr[a]) * w;
                const float ar = (gr * qr - gi * qi) * w, ai = (gr * qi + gi * qr) * w;
                const float bq = br * K.qre - bi * K.qim;   // Re[beta G Q / r^2]
                const float bh = br * K.hre - bi * K.him;   // Re[beta H]
                const float ah = ar * K.hre - ai * K.him;   // Re[alpha H]
                const float ag = ar * K.gre - ai * K.gim;   // Re[alpha G]
                const float cr = fmaf(bq, rn, -jac * ah);
                const float gx = fmaf(cr, rx, bh * nx), gyv = fmaf(cr, ry, bh * ny),
                            gz = fmaf(cr, rz, bh * nz);
                gya[0][0] = fmaf(b0, gx, gya[0][0]);
                gya[0][1] = fmaf(b0, gyv, gya[0][1]);
                gya[0][2] = fmaf(b0, gz, gya[0][2]);
                gya[1][0] = fmaf(xi, gx, gya[1][0]);
                gya[1][1] = fmaf(xi, gyv, gya[1][1]);
                gya[1][2] = fmaf(xi, gz, gya[1][2]);
                gya[2][0] = fmaf(eta, gx, gya[2][0]);
                gya[2][1] = fmaf(eta, gyv, gya[2][1]);
                gya[2][2] = fmaf(eta, gz, gya[2][2]);
                // d/dN: Re[beta H] r - Re[alpha G] N^
                gna[0] = fmaf(bh, rx, fmaf(-ag, nhx, gna[0]));
                gna[1] = fmaf(bh, ry, fmaf(-ag, nhy, gna[1]));
                gna[2] = fmaf(bh, rz, fmaf(-ag, nhz, gna[2]));
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cj[j][0] += c_re[j];
            cj[j][1] += c_im[j];
            gn[j] += gna[j];
#pragma unroll
            for (int d = 0; d < 3; ++d) gy[j][d] += gya[j][d];
        }
        ftd[0] += f_re;
        ftd[1] += f_im;
    }
    if (!live || o_begin >= o_end) return;
    double *cb = c_out + (size_t)b * c_stride * 2;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        atomicAdd(cb + 2 * dof[j], cj[j][0]);
        atomicAdd(cb + 2 * dof[j] + 1, cj[j][1]);
    }
    atomicAdd(ft_out + ((size_t)b * T + orig) * 2, ftd[0]);
    atomicAdd(ft_out + ((size_t)b * T + orig) * 2 + 1, ftd[1]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int v = ti[(size_t)(3 + j) * T + tdd];
        atomic_add3(gV, nv, v, gy[j][0], gy[j][1], gy[j][2]);
    }
    atomic_add3(gN, T, orig, gn[0], gn[1], gn[2]);
}

// Right-hand side of the adjoint solve: e = conj(c) + R^H conj(ft), written
// as interleaved float (the LU solve input) and double (the refinement
// target), zero beyond n.  One thread per (dof, frequency); the Robin term
// is scattered per triangle by k_adjoint_robin_rhs beforehand into c.
__global__ void k_adjoint_rhs(const double *__restrict__ c, int64_t ld, int n, int N,
                              float *__restrict__ e32, double *__restrict__ e64) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= N) return;
    const size_t idx = ((size_t)b * ld + i) * 2;
    const double re = i < n ? c[idx] : 0.0, im = i < n ? -c[idx + 1] : 0.0;
    e32[idx] = (float)re;
    e32[idx + 1] = (float)im;
    e64[idx] = re;
    e64[idx + 1] = im;
}

// c_j += conj(i k beta_t / 3) conj(ft_t) conj-folded: the kernel adds to c
// (which is conjugated afterwards) the value (i k beta_t / 3) ft_t for each
// dof j of triangle t, so that conj(c) gains conj(i k beta / 3) conj(ft).
__global__ void k_adjoint_robin_rhs(const int32_t *__restrict__ l2g, int T,
                                    const float *__restrict__ k_real,
                                    const float *__restrict__ k_imag,
                                    const float *__restrict__ beta, const double *__restrict__ ft,
                                    double *__restrict__ c, int64_t ld) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (t >= T) return;
    const double br = beta[((size_t)b * T + t) * 2], bi = beta[((size_t)b * T + t) * 2 + 1];
    if (br == 0.0 && bi == 0.0) return;
    const double ikr = k_imag ? -(double)k_imag[b] : 0.0, iki = k_real[b];
    const double cr = (ikr * br - iki * bi) / 3.0, ci = (ikr * bi + iki * br) / 3.0;
    const double fr = ft[((size_t)b * T + t) * 2], fi = ft[((size_t)b * T + t) * 2 + 1];
    const double vr = cr * fr - ci * fi, vi = cr * fi + ci * fr;
    for (int l = 0; l < 3; ++l) {
        const int d = l2g[3 * t + l];
        atomicAdd(c + ((size_t)b * ld + d) * 2, vr);
        atomicAdd(c + ((size_t)b * ld + d) * 2 + 1, vi);
    }
}

// r = e - A^H x in FP64 (A FP32 column-major interleaved), per-system
// max-norms of r and e into norms[2 b], norms[2 b + 1].  One wave per
// column j: (A^H x)_j = sum_i conj(a_ij) x_i.
__global__ __launch_bounds__(256) void k_residual_h_f64(const float *__restrict__ A, int64_t lda,
                                                        int64_t a_stride, int n,
                                                        const double *__restrict__ e,
                                                        const double *__restrict__ x, int64_t ldx,
                                                        double *__restrict__ r,
                                                        double *__restrict__ norms, int vpm) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int b = blockIdx.y;  // vector; its matrix is b / vpm
    if (j >= n) return;
    const float *col = A + ((size_t)(b / vpm) * a_stride + (size_t)j * lda) * 2;
    const double *xb = x + (size_t)b * ldx * 2;
    double sr = 0.0, si = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double ar = col[2 * i], ai = col[2 * i + 1];
        const double xr = xb[2 * i], xi = xb[2 * i + 1];
        // conj(a) x
        sr = fma(ar, xr, fma(ai, xi, sr));
        si = fma(ar, xi, fma(-ai, xr, si));
    }
    for (int o = 32; o >= 1; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
    }
    if (lane == 0) {
        const size_t idx = ((size_t)b * ldx + j) * 2;
        const double rr = e[idx] - sr, ri = e[idx + 1] - si;
        r[idx] = rr;
        r[idx + 1] = ri;
        const double rn = fmax(fabs(rr), fabs(ri)), en = fmax(fabs(e[idx]), fabs(e[idx + 1]));
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b),
                  static_cast<unsigned long long>(__double_as_longlong(rn)));
        atomicMax(reinterpret_cast<unsigned long long *>(norms + 2 * b + 1),
                  static_cast<unsigned long long>(__double_as_longlong(en)));
    }
}

// ---------------------------------------------------------------------------
// Regular-pair adjoint.  A wave owns 64 trial triangles (lanes, device
// order) and walks a range of test triangles (uniform).  For each test t,
// image m and point pair (a, b), in the trial frame r = y_b - S_m x_a with
// N = N_s:
//   phi = Re[L_a (qt_s |N_s| G - P_b H (r . N))],  L_a = rw w_a |N_t| sum_i
//         conj(lambda_{row_i}) phi_i(a),  times w_b
// Trial-side sums stay in registers (FP64 across tests), test-side sums are
// reduced over the wave per test.
// ---------------------------------------------------------------------------
constexpr int kAdjTestsPerWave = 96;

// Two waves per SIMD (at most 256 VGPRs; unbounded the kernel took 282 and
// ran one wave per SIMD with its transcendental chains exposed).
template <bool kComplexK>
__global__ __launch_bounds__(256, 2) void k_adjoint_regular(
    DeviceGeometry g, const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ lam, const float *__restrict__ p, int64_t ld,
    const float *__restrict__ qt, double *__restrict__ gV, double *__restrict__ gN,
    double *__restrict__ sig) {
    const int T = g.n_triangles, nv = g.n_vertices;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n_chunks = (T + 63) / 64;
    const int chunk = wave % n_chunks;
    const int split = wave / n_chunks;
    const int t_begin = split * kAdjTestsPerWave;
    if (t_begin >= T) return;
    const int t_end = min(T, t_begin + kAdjTestsPerWave);
    const int b = blockIdx.y;
    const int sd = chunk * 64 + lane;
    const bool live = sd < T;
    const int sdd = live ? sd : T - 1;
    const float *tfs = g.tri_f_soa;
    const int32_t *tis = g.tri_i_soa;
    // trial points as vertex 0 plus offsets X e1 + Y e2 (edges of exact
    // float vertices): differences of nearby points then round relative to
    // their own size, not to the coordinates'
    float v0s[3], ys[18];
    {
        float e1[3], e2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v0s[c] = tfs[(size_t)(23 + c) * T + sdd];
            e1[c] = tfs[(size_t)(26 + c) * T + sdd] - v0s[c];
            e2[c] = tfs[(size_t)(29 + c) * T + sdd] - v0s[c];
        }
#pragma unroll
        for (int b = 0; b < 6; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                ys[3 * b + c] = fmaf((float)kRule6X[b], e1[c], (float)kRule6Y[b] * e2[c]);
    }
    const float js = tfs[(size_t)22 * T + sdd];
    const float nhx = tfs[(size_t)18 * T + sdd], nhy = tfs[(size_t)19 * T + sdd],
                nhz = tfs[(size_t)20 * T + sdd];
    const float nx = nhx * js, ny = nhy * js, nz = nhz * js;
    int skey[kMaxMasks][3];
#pragma unroll
    for (int m = 0; m < kMaxMasks; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            skey[m][j] = m < g.n_masks ? tis[(size_t)(6 + 3 * m + j) * T + sdd] : -2;
    const int s_orig = tis[(size_t)19 * T + sdd];
    const float *pb = p + (size_t)b * ld * 2;
    const float *lb = lam + (size_t)b * ld * 2;
    float Pr[6], Pi[6];
    {
        float pr[3], pim[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const int d = tis[(size_t)l * T + sdd];
            pr[l] = pb[2 * d];
            pim[l] = pb[2 * d + 1];
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
            Pr[a] = b0 * pr[0] + xi * pr[1] + eta * pr[2];
            Pi[a] = b0 * pim[0] + xi * pim[1] + eta * pim[2];
        }
    }
    const float qr = qt[((size_t)b * T + s_orig) * 2], qi = qt[((size_t)b * T + s_orig) * 2 + 1];
    const float k = k_real[b], kim = kComplexK ? k_imag[b] : 0.0f;
    const double krev = (double)k * kInv2PiD;
    const float rw = g.row_weight;

    double dy[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dn[3] = {0, 0, 0}, dnh = 0.0;
    double ds[2] = {0, 0};
    const float *tfa = g.tri_f_aos;
    const int32_t *tia = g.tri_i_aos;
    for (int t = t_begin; t < t_end; ++t) {
        const float *tf = tfa + (size_t)t * kTriF;
        const int32_t *ti = tia + (size_t)t * kTriI;
        const float jt = tf[22];
        // test points as vertex 0 plus X_a e1 + Y_a e2, the offsets formed
        // where they are used (registers: the kernel's occupancy is set by
        // its VGPR count)
        float v0t[3], te1[3], te2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v0t[c] = tf[23 + c];
            te1[c] = tf[26 + c] - v0t[c];
            te2[c] = tf[29 + c] - v0t[c];
        }
        // L_a = rw w_a |N_t| sum_i conj(lambda_i) phi_i(a)
        float lr[3], li[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            lr[i] = lb[2 * ti[i]];
            li[i] = -lb[2 * ti[i] + 1];
        }
        float Lr[6], Li[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
            const float sc = rw * (float)kRule6W[a] * jt;
            Lr[a] = sc * (b0 * lr[0] + xi * lr[1] + eta * lr[2]);
            Li[a] = sc * (b0 * li[0] + xi * li[1] + eta * li[2]);
        }
        float ty[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, tn[3] = {0, 0, 0}, tnh = 0.0f;
        float tsr = 0.0f, tsi = 0.0f;
        float xg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // test vertex gradients
        float gjt = 0.0f;
        for (int m = 0; m < g.n_masks; ++m) {
            bool touch = false;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) touch |= ti[6 + i] == skey[m][j];
            if (touch || !live) continue;
            const int mask = g.masks[m];
            const float sx = msign(mask, 1), sy = msign(mask, 2), sz = msign(mask, 4);
            // y_b - S_m x_a = (v0_s - S_m v0_t) + (o_b - S_m o_a)
            const float ax = v0s[0] - sx * v0t[0], ay = v0s[1] - sy * v0t[1],
                        az = v0s[2] - sz * v0t[2];
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const float xoa = fmaf((float)kRule6X[a], te1[0], (float)kRule6Y[a] * te2[0]);
                const float yoa = fmaf((float)kRule6X[a], te1[1], (float)kRule6Y[a] * te2[1]);
                const float zoa = fmaf((float)kRule6X[a], te1[2], (float)kRule6Y[a] * te2[2]);
                const float bx = ax - sx * xoa, by = ay - sy * yoa, bz = az - sz * zoa;
                const float alr = Lr[a] * qr - Li[a] * qi, ali = Lr[a] * qi + Li[a] * qr;
                float gxa = 0.0f, gya = 0.0f, gza = 0.0f;
                float gsr = 0.0f, gsi = 0.0f;  // sum_b w_b G
                float ja = 0.0f;               // sum_b w_b (-Re[beta H] (r . N))
#pragma unroll
                for (int bb = 0; bb < 6; ++bb) {
                    const float wb = (float)kRule6W[bb];
                    const float rx = bx + ys[3 * bb], ry = by + ys[3 * bb + 1], rz = bz + ys[3 * bb + 2];
                    const AdjK K = adj_kernel_p<kComplexK>(rx, ry, rz, k, krev, kim);
                    const float rn = fmaf(rx, nx, fmaf(ry, ny, rz * nz));
                    const float btr = (Lr[a] * Pr[bb] - Li[a] * Pi[bb]) * wb;
                    const float bti = (Lr[a] * Pi[bb] + Li[a] * Pr[bb]) * wb;
                    const float a1 = (alr * K.hre - ali * K.him) * wb;  // Re[alpha H] w_b
                    const float a2 = btr * K.qre - bti * K.qim;         // Re[beta G Q / r^2]
                    const float a3 = btr * K.hre - bti * K.him;         // Re[beta H]
                    const float cr = fmaf(js, a1, -a2 * rn);
                    const float gx = fmaf(cr, rx, -a3 * nx), gy = fmaf(cr, ry, -a3 * ny),
                                gz = fmaf(cr, rz, -a3 * nz);
                    gxa += gx;
                    gya += gy;
                    gza += gz;
                    const float xi = (float)kRule6X[bb], eta = (float)kRule6Y[bb],
                                b0 = 1.0f - xi - eta;
                    ty[0][0] = fmaf(b0, gx, ty[0][0]);
                    ty[0][1] = fmaf(b0, gy, ty[0][1]);
                    ty[0][2] = fmaf(b0, gz, ty[0][2]);
                    ty[1][0] = fmaf(xi, gx, ty[1][0]);
                    ty[1][1] = fmaf(xi, gy, ty[1][1]);
                    ty[1][2] = fmaf(xi, gz, ty[1][2]);
                    ty[2][0] = fmaf(eta, gx, ty[2][0]);
                    ty[2][1] = fmaf(eta, gy, ty[2][1]);
                    ty[2][2] = fmaf(eta, gz, ty[2][2]);
                    // d/dN_s: Re[alpha G] N^ - Re[beta H] r
                    tn[0] = fmaf(-a3, rx, tn[0]);
                    tn[1] = fmaf(-a3, ry, tn[1]);
                    tn[2] = fmaf(-a3, rz, tn[2]);
                    gsr = fmaf(wb, K.gre, gsr);
                    gsi = fmaf(wb, K.gim, gsi);
                    ja = fmaf(-a3, rn, ja);
                }
                // Re[alpha sum_b w_b G]
                const float ag = alr * gsr - ali * gsi;
                tnh += ag;
                gjt += fmaf(js, ag, ja);
                // sigma: L_a sum_b w_b G (times |N_s| at the end)
                tsr += Lr[a] * gsr - Li[a] * gsi;
                tsi += Lr[a] * gsi + Li[a] * gsr;
                // test points move with -S_m
                const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a], b0 = 1.0f - xi - eta;
                const float mx = -sx * gxa, my = -sy * gya, mz = -sz * gza;
                xg[0][0] = fmaf(b0, mx, xg[0][0]);
                xg[0][1] = fmaf(b0, my, xg[0][1]);
                xg[0][2] = fmaf(b0, mz, xg[0][2]);
                xg[1][0] = fmaf(xi, mx, xg[1][0]);
                xg[1][1] = fmaf(xi, my, xg[1][1]);
                xg[1][2] = fmaf(xi, mz, xg[1][2]);
                xg[2][0] = fmaf(eta, mx, xg[2][0]);
                xg[2][1] = fmaf(eta, my, xg[2][1]);
                xg[2][2] = fmaf(eta, mz, xg[2][2]);
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            dn[j] += tn[j];
#pragma unroll
            for (int d = 0; d < 3; ++d) dy[j][d] += ty[j][d];
        }
        dnh += tnh;
        ds[0] += tsr;
        ds[1] += tsi;
        // wave reduction of the test-side sums (9 point gradients + |N_t|)
        float red[10];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int d = 0; d < 3; ++d) red[3 * j + d] = xg[j][d];
        red[9] = gjt / jt;
#pragma unroll
        for (int i = 0; i < 10; ++i)
            for (int o = 32; o >= 1; o >>= 1) red[i] += __shfl_xor(red[i], o);
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < 3; ++j) atomic_add3(gV, nv, ti[3 + j], red[3 * j], red[3 * j + 1], red[3 * j + 2]);
            // |N_t| enters through N^_t
            atomic_add3(gN, T, ti[19], (double)red[9] * tf[18], (double)red[9] * tf[19],
                        (double)red[9] * tf[20]);
        }
    }
    if (!live) return;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int v = tis[(size_t)(3 + j) * T + sdd];
        atomic_add3(gV, nv, v, dy[j][0], dy[j][1], dy[j][2]);
    }
    atomic_add3(gN, T, s_orig, dn[0] + dnh * nhx, dn[1] + dnh * nhy, dn[2] + dnh * nhz);
    atomicAdd(sig + ((size_t)b * T + s_orig) * 2, ds[0] * js);
    atomicAdd(sig + ((size_t)b * T + s_orig) * 2 + 1, ds[1] * js);
}

// ---------------------------------------------------------------------------
// Singular-class adjoint: one lane per class, the class's Sauter-Schwab
// rule in runs of kSgRun points (FP32) carried in FP64.  Image frame: r =
// y - x with both triangles mirrored (test by tm, trial by sm), n^ the
// mirrored unit trial normal.  With wc = |N_t| |N_s| count (the class
// scale) the class adds wc U, U = sum_q Re[L_q (qt G - P_q H (r . n^))],
// L_q = sum_a conj(lambda_a) w rho_a.  Then
//   d/dx_corner = -wc rho_a grad_r,  d/dy_corner = wc rho'_b grad_r
//   d/d|N_t|    = wc U / |N_t|
//   d/dN_img    = (wc / |N_s|) sum_q (Re[L qt G] n^ - Re[L P H] r)
// and the corner / normal gradients go back to the caller frame by S_tm,
// S_sm.
// ---------------------------------------------------------------------------
template <bool kComplexK>
__global__ __launch_bounds__(256) void k_adjoint_singular(
    DeviceGeometry g, const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ lam, const float *__restrict__ p, int64_t ld,
    const float *__restrict__ qt, double *__restrict__ gV, double *__restrict__ gN,
    double *__restrict__ sig) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int w1 = g.sg_waves[0], w2 = w1 + g.sg_waves[1], w3 = w2 + g.sg_waves[2];
    if (w >= w3) return;
    const int kind = w < w1 ? 0 : (w < w2 ? 1 : 2);
    const int np = g.sg_np;
    const int c = w * 64 + lane;
    const int orig = g.sg_i[(size_t)6 * np + c];
    const bool live = orig >= 0;
    const int T = g.n_triangles, nv = g.n_vertices;
    const int b = blockIdx.y;
    const float *__restrict__ F = g.sg_f;
    float gf[18];
#pragma unroll
    for (int j = 0; j < 18; ++j) gf[j] = F[(size_t)j * np + c];
    const double wc = F[(size_t)18 * np + c];
    int rows[3], cols[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rows[a] = live ? g.sg_i[(size_t)a * np + c] : 0;
        cols[a] = live ? g.sg_i[(size_t)(3 + a) * np + c] : 0;
    }
    const int so = live ? orig : 0;
    const float *pb = p + (size_t)b * ld * 2;
    const float *lb = lam + (size_t)b * ld * 2;
    float lr[3], li[3], pr[3], pim[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lr[a] = lb[2 * rows[a]];
        li[a] = -lb[2 * rows[a] + 1];
        pr[a] = pb[2 * cols[a]];
        pim[a] = pb[2 * cols[a] + 1];
    }
    const float qr = qt[((size_t)b * T + so) * 2], qi = qt[((size_t)b * T + so) * 2 + 1];
    const float k = k_real[b], krev = k * kInv2Pi, kim = kComplexK ? k_imag[b] : 0.0f;
    const float nhx = gf[15], nhy = gf[16], nhz = gf[17];

    // A corner shared by the test and trial triangles (the same caller vertex
    // at the same rule-local index on both sides) moves both points of every
    // rule pair: along the components where the two image signs agree its
    // derivative is wc sum_q (rho'_a(y) - rho_a(x)) grad_r, accumulated with
    // the exact weight difference (linear in the stored (x'-x, y'-y)) instead
    // of as two large cancelling sums.
    const int packed0 = live ? g.sg_adj[(size_t)6 * np + c] : 0;
    const int tm0 = packed0 & 15, sm0 = packed0 >> 4;
    bool comb[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const bool same_v =
            live && g.sg_adj[(size_t)j * np + c] == g.sg_adj[(size_t)(3 + j) * np + c];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            comb[j][d] = same_v && (((tm0 >> d) & 1) == ((sm0 >> d) & 1));
    }
    const float *__restrict__ rp = g.sg_rule + (size_t)g.sg_rule_offset[kind] * kSgRule;
    const int cnt = g.sg_rule_count[kind];
    double dx[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double dyv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double dn[3] = {0, 0, 0}, da4 = 0.0, du = 0.0, dsr = 0.0, dsi = 0.0;
    for (int p0 = 0; p0 < cnt; p0 += kSgRun) {
        float ax[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        float ay[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        float an[3] = {0, 0, 0}, aa4 = 0.0f, au = 0.0f, asr = 0.0f, asi = 0.0f;
        for (int pp = p0; pp < p0 + kSgRun; ++pp) {
            const float *pt = rp + (size_t)pp * kSgRule;
            const float ddx = pt[0], ddy = pt[1], xs = pt[2], ys = pt[3];
            const float rx = fmaf(ys, gf[12], fmaf(xs, gf[9], fmaf(ddy, gf[6], fmaf(ddx, gf[3], gf[0]))));
            const float ry = fmaf(ys, gf[13], fmaf(xs, gf[10], fmaf(ddy, gf[7], fmaf(ddx, gf[4], gf[1]))));
            const float rz = fmaf(ys, gf[14], fmaf(xs, gf[11], fmaf(ddy, gf[8], fmaf(ddx, gf[5], gf[2]))));
            const AdjK K = adj_kernel<kComplexK>(rx, ry, rz, k, krev, kim);
            const float w0 = pt[13], w1v = pt[14], w2v = pt[15];
            const float wq = w0 + w1v + w2v;
            const float Lr = w0 * lr[0] + w1v * lr[1] + w2v * lr[2];
            const float Li = w0 * li[0] + w1v * li[1] + w2v * li[2];
            const float s0 = 1.0f - xs - ys;
            const float Pr = s0 * pr[0] + xs * pr[1] + ys * pr[2];
            const float Pi = s0 * pim[0] + xs * pim[1] + ys * pim[2];
            const float rn = fmaf(rx, nhx, fmaf(ry, nhy, rz * nhz));
            const float alr = Lr * qr - Li * qi, ali = Lr * qi + Li * qr;
            const float btr = Lr * Pr - Li * Pi, bti = Lr * Pi + Li * Pr;
            const float a1 = alr * K.hre - ali * K.him;
            const float a2 = btr * K.qre - bti * K.qim;
            const float a3 = btr * K.hre - bti * K.him;
            const float a4 = alr * K.gre - ali * K.gim;
            const float cr = fmaf(-a2, rn, a1);
            const float gx = fmaf(cr, rx, -a3 * nhx), gy = fmaf(cr, ry, -a3 * nhy),
                        gz = fmaf(cr, rz, -a3 * nhz);
            const float iw = 1.0f / wq;
            const float ra[3] = {w0 * iw, w1v * iw, w2v * iw};
            const float sb[3] = {s0, xs, ys};
            // rho'_a(x', y') - rho_a(x, y) of the rule-local basis
            const float dr[3] = {-(ddx + ddy), ddx, ddy};
            const float gg[3] = {gx, gy, gz};
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    ax[j][d] = fmaf(comb[j][d] ? dr[j] : -ra[j], gg[d], ax[j][d]);
                    ay[j][d] = fmaf(comb[j][d] ? 0.0f : sb[j], gg[d], ay[j][d]);
                }
            an[0] = fmaf(-a3, rx, an[0]);
            an[1] = fmaf(-a3, ry, an[1]);
            an[2] = fmaf(-a3, rz, an[2]);
            aa4 += a4;
            au += fmaf(-a3, rn, a4);
            asr += Lr * K.gre - Li * K.gim;
            asi += Lr * K.gim + Li * K.gre;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            dn[j] += an[j];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dx[j][d] += ax[j][d];
                dyv[j][d] += ay[j][d];
            }
        }
        da4 += aa4;
        du += au;
        dsr += asr;
        dsi += asi;
    }
    if (!live) return;
    const int packed = g.sg_adj[(size_t)6 * np + c];
    const int tm = packed & 15, sm = packed >> 4;
    const int t_orig = g.sg_adj[(size_t)7 * np + c];
    const double ts[3] = {msign(tm, 1), msign(tm, 2), msign(tm, 4)};
    const double ss[3] = {msign(sm, 1), msign(sm, 2), msign(sm, 4)};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int vt = g.sg_adj[(size_t)j * np + c], vs = g.sg_adj[(size_t)(3 + j) * np + c];
        atomic_add3(gV, nv, vt, wc * ts[0] * dx[j][0], wc * ts[1] * dx[j][1], wc * ts[2] * dx[j][2]);
        atomic_add3(gV, nv, vs, wc * ss[0] * dyv[j][0], wc * ss[1] * dyv[j][1], wc * ss[2] * dyv[j][2]);
    }
    const double jt = g.jac_c[t_orig], js = g.jac_c[so];
    const double ut = wc * du / jt;
    atomic_add3(gN, T, t_orig, ut * g.normals[t_orig], ut * g.normals[T + t_orig],
                ut * g.normals[2 * T + t_orig]);
    const double kn = wc / js;
    atomic_add3(gN, T, so, ss[0] * kn * (da4 * nhx + dn[0]), ss[1] * kn * (da4 * nhy + dn[1]),
                ss[2] * kn * (da4 * nhz + dn[2]));
    atomicAdd(sig + ((size_t)b * T + so) * 2, wc * dsr);
    atomicAdd(sig + ((size_t)b * T + so) * 2 + 1, wc * dsi);
}

// ---------------------------------------------------------------------------
// Per caller triangle: the mass term's |N_t| derivative (A carries
// -rw |N_t| / 4 Mhat on the triangle's dofs: d/d|N_t| of Re(lambda^H (-A) p)
// is rw / 4 Re(lambda^H Mhat p), summed over frequencies), then the normal
// gradient d/dN_t -> vertex gradients with N = s (v1 - v0) x (v2 - v0),
// s = +-1 the orientation of the stored normal.
// ---------------------------------------------------------------------------
__global__ void k_adjoint_normals(DeviceGeometry g, int n_freq, const float *__restrict__ lam,
                                  const float *__restrict__ p, int64_t ld,
                                  const double *__restrict__ gN, double *__restrict__ gV) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int T = g.n_triangles, nv = g.n_vertices;
    if (t >= T) return;
    int d[3], v[3];
    for (int l = 0; l < 3; ++l) {
        d[l] = g.l2g[3 * t + l];
        v[l] = g.tri_v[l * T + t];
    }
    double mass = 0.0;
    for (int b = 0; b < n_freq; ++b) {
        const float *lb = lam + (size_t)b * ld * 2, *pb = p + (size_t)b * ld * 2;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double m = i == j ? 1.0 / 6.0 : 1.0 / 12.0;
                // Re(conj(lambda_i) p_j)
                mass += m * ((double)lb[2 * d[i]] * pb[2 * d[j]] +
                             (double)lb[2 * d[i] + 1] * pb[2 * d[j] + 1]);
            }
    }
    const double nh[3] = {g.normals[t], g.normals[T + t], g.normals[2 * T + t]};
    const double gm = 0.25 * (double)g.row_weight * mass;
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = gN[(size_t)k * T + t] + gm * nh[k];
    double x[3][3];
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 3; ++k) x[l][k] = g.vtx[k * nv + v[l]];
    double e1[3], e2[3], e12[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = x[1][k] - x[0][k];
        e2[k] = x[2][k] - x[0][k];
        e12[k] = x[1][k] - x[2][k];
    }
    auto cross = [](const double *a, const double *b, double *o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double nrm[3];
    cross(e1, e2, nrm);
    const double sgn = (nrm[0] * nh[0] + nrm[1] * nh[1] + nrm[2] * nh[2]) >= 0.0 ? 1.0 : -1.0;
    double g0[3], g1[3], g2[3];
    cross(e12, c, g0);
    cross(e2, c, g1);
    cross(c, e1, g2);
    atomic_add3(gV, nv, v[0], sgn * g0[0], sgn * g0[1], sgn * g0[2]);
    atomic_add3(gV, nv, v[1], sgn * g1[0], sgn * g1[1], sgn * g1[2]);
    atomic_add3(gV, nv, v[2], sgn * g2[0], sgn * g2[1], sgn * g2[2]);
}

// grad_q [F][T] (interleaved complex float) = conj(ft + sigma).
__global__ void k_adjoint_grad_q(int n, const double *__restrict__ ft,
                                 const double *__restrict__ sig, float *__restrict__ gq) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    gq[2 * i] = (float)(ft[2 * i] + sig[2 * i]);
    gq[2 * i + 1] = (float)(-(ft[2 * i + 1] + sig[2 * i + 1]));
}

}  // namespace

void launch_field_adjoint(const DeviceGeometry &g, int n_batch, const float *field_k,
                          const float *p, int64_t p_stride, const float *qt, int n_obs,
                          const float *obs, const float *grad_field, double *c, int64_t c_stride,
                          double *ft, double *gV, double *gN, hipStream_t stream) {
    if (n_obs == 0 || n_batch <= 0) return;
    dim3 grid((g.n_triangles + 255) / 256, (n_obs + kFieldAdjChunk - 1) / kFieldAdjChunk, n_batch);
    hipLaunchKernelGGL(k_field_adjoint, grid, dim3(256), 0, stream, g, field_k, p, p_stride, qt,
                       n_obs, obs, grad_field, c, c_stride, ft, gV, gN);
}

void launch_adjoint_rhs(const DeviceGeometry &g, int n_batch, const float *k_real,
                        const float *k_imag, const float *beta, const double *ft, double *c,
                        int64_t ld, int n, int N, float *e32, double *e64, hipStream_t stream) {
    if (beta)
        hipLaunchKernelGGL(k_adjoint_robin_rhs, dim3((g.n_triangles + 255) / 256, n_batch),
                           dim3(256), 0, stream, g.l2g, g.n_triangles, k_real, k_imag, beta, ft, c,
                           ld);
    hipLaunchKernelGGL(k_adjoint_rhs, dim3((N + 255) / 256, n_batch), dim3(256), 0, stream, c, ld,
                       n, N, e32, e64);
}

void launch_residual_h_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                           const double *e, const double *x, int64_t ldx, double *r,
                           double *norms, hipStream_t stream, int vpm) {
    launch_fill_zero_f64(norms, (size_t)2 * B, stream);
    hipLaunchKernelGGL(k_residual_h_f64, dim3((n + 3) / 4, B), dim3(256), 0, stream, A, lda,
                       a_stride, n, e, x, ldx, r, norms, vpm);
}

void launch_assembly_adjoint(const DeviceGeometry &g, int n_batch, const float *k_real,
                             const float *k_imag, const float *lam, const float *p, int64_t ld,
                             const float *qt, double *gV, double *gN, double *sig,
                             hipStream_t stream) {
    if (n_batch <= 0) return;
    const int T = g.n_triangles;
    const int waves = ((T + 63) / 64) * ((T + kAdjTestsPerWave - 1) / kAdjTestsPerWave);
    dim3 grid((waves + 3) / 4, n_batch);
    if (k_imag)
        hipLaunchKernelGGL((k_adjoint_regular<true>), grid, dim3(256), 0, stream, g, k_real, k_imag,
                           lam, p, ld, qt, gV, gN, sig);
    else
        hipLaunchKernelGGL((k_adjoint_regular<false>), grid, dim3(256), 0, stream, g, k_real,
                           k_imag, lam, p, ld, qt, gV, gN, sig);
    const int sw = g.sg_waves[0] + g.sg_waves[1] + g.sg_waves[2];
    if (sw > 0) {
        dim3 gs((sw + 3) / 4, n_batch);
        if (k_imag)
            hipLaunchKernelGGL((k_adjoint_singular<true>), gs, dim3(256), 0, stream, g, k_real,
                               k_imag, lam, p, ld, qt, gV, gN, sig);
        else
            hipLaunchKernelGGL((k_adjoint_singular<false>), gs, dim3(256), 0, stream, g, k_real,
                               k_imag, lam, p, ld, qt, gV, gN, sig);
    }
}

void launch_adjoint_finish(const DeviceGeometry &g, int n_batch, const float *lam, const float *p,
                           int64_t ld, const double *gN, double *gV, const double *ft,
                           const double *sig, float *grad_q, hipStream_t stream) {
    hipLaunchKernelGGL(k_adjoint_normals, dim3((g.n_triangles + 255) / 256), dim3(256), 0, stream,
                       g, n_batch, lam, p, ld, gN, gV);
    if (grad_q) {
        const int n = n_batch * g.n_triangles;
        hipLaunchKernelGGL(k_adjoint_grad_q, dim3((n + 255) / 256), dim3(256), 0, stream, n, ft,
                           sig, grad_q);
    }
}

}  // namespace hbem
