// gfx950 kernels of the assembly and field stages of libhbem.
//
//   k_assemble_regular  -- every non-touching (test, trial, image) triangle
//                          pair with the 6x6-point rule, fused double-layer
//                          3x3 block + single-layer right-hand side
//                          (reference: assemble_matrix_pair_atomic and
//                          assemble_rhs_source_regular, main.swift:3698-3736
//                          and 3993-4133).
//   k_assemble_singular -- touching pair classes with the Sauter-Schwab rules
//                          plus the -M/2 mass term (reference:
//                          duffy_delta_blocks main.swift:4246-4431 folded by
//                          reduceDuffyDeltaBlocks main.swift:6498-6561).
//   k_assemble_near     -- optional near-field deltas (subdivided minus
//                          6x6 rule) of close non-touching pairs
//                          (applyNearFieldCorrectionsCPU, main.swift:3072-3175).
//   k_field             -- exterior potential (evaluate_field_regular,
//                          main.swift:4135-4203).
//   k_tag_average       -- area-weighted tag means
//                          (averageSurfacePressureForTag, main.swift:807-841).
//
// Matrices are column-major interleaved complex (re, im) float with leading
// dimension lda; vectors are interleaved complex.
//
// Kernel of every stage, for r = y - x (trial minus test / source minus
// observer), complex k = k_re + i k_im:
//   G      = e^{i k_re r} e^{-k_im r} / (4 pi r)
//   dG/dn  = G (i k_re - k_im - 1/r) (r . n_y) / r
// The phase is range-reduced in revolutions (t = k_re r / 2pi, t -= rint(t))
// before the hardware sin/cos, which take revolutions.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hbem_internal.h"

namespace hbem {
namespace {

constexpr float kInv4Pi = 0.07957747154594767f;
constexpr float kInv2Pi = 0.15915494309189535f;

struct GD {
    float gre, gim, dre, dim;
};

// G and dG/dn_y for d = y - x and normal n (FP32).  Returns zeros for
// r^2 <= 1e-14 (the reference GPU guard, main.swift:3443-3480).
template <bool kComplexK>
__device__ __forceinline__ GD kernel_gd(float dx, float dy, float dz, float nx, float ny,
                                        float nz, float k, float krev, float kim) {
    const float r2raw = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool ok = r2raw > 1.0e-14f;
    const float r2 = ok ? r2raw : 1.0e-14f;  // keeps every lane finite
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float t = r * krev;
    t = t - __builtin_rintf(t);
    const float s = __builtin_amdgcn_sinf(t);
    const float c = __builtin_amdgcn_cosf(t);
    float scale = kInv4Pi * rinv;
    if (kComplexK) scale *= __expf(-kim * r);
    scale = ok ? scale : 0.0f;
    const float gre = c * scale;
    const float gim = s * scale;
    const float proj = fmaf(dx, nx, fmaf(dy, ny, dz * nz)) * rinv;
    const float fre = kComplexK ? (-rinv - kim) : -rinv;
    GD o;
    o.gre = gre;
    o.gim = gim;
    o.dre = (gre * fre - gim * k) * proj;
    o.dim = (gre * k + gim * fre) * proj;
    return o;
}

__device__ __forceinline__ float mask_sign(int mask, int bit) { return (mask & bit) ? -1.0f : 1.0f; }

constexpr double kInv2PiD = 0.15915494309189533577;

// G and dG/dn_y at an observation point for d = y - x given in FP64: the
// distance and the phase k r (in revolutions, reduced to [-1/2, 1/2]) are
// formed in FP64, so a field point metres away keeps the phase accuracy of
// the quadrature points; sin, cos and the amplitudes are FP32.
__device__ __forceinline__ GD kernel_gd_far(double dx, double dy, double dz, float nx, float ny,
                                            float nz, float k, double krev) {
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const bool ok = r2 > 1.0e-14;
    const float r2f = ok ? (float)r2 : 1.0f;
    const float rinv32 = __builtin_amdgcn_rsqf(r2f);
    const float r32 = r2f * rinv32;
    // one Newton step from the FP32 root: relative error ~1e-14
    const double r = (double)r32 + fma(-(double)r32, (double)r32, r2) * (double)(0.5f * rinv32);
    double t = r * krev;
    t -= rint(t);
    const float tf = (float)t;
    const float s = __builtin_amdgcn_sinf(tf);
    const float c = __builtin_amdgcn_cosf(tf);
    const float rinv = ok ? rinv32 : 0.0f;
    const float scale = kInv4Pi * rinv;
    const float gre = c * scale, gim = s * scale;
    const float proj = fmaf((float)dx, nx, fmaf((float)dy, ny, (float)dz * nz)) * rinv;
    const float fre = -rinv;
    GD o;
    o.gre = gre;
    o.gim = gim;
    o.dre = (gre * fre - gim * k) * proj;
    o.dim = (gre * k + gim * fre) * proj;
    return o;
}

// ---------------------------------------------------------------------------
// Regular pairs.
// One block = one unordered pair {T, S} of tiles of up to 64 triangles, for
// one frequency (grid.y).  Lanes own the triangles s of S (records in
// registers); each wave walks 16 triangles t of T, whose records are
// wave-uniform (scalar loads).  The quadrature kernel is reciprocal: for an
// image mask m the point sets of (test t, trial s^m) and (test s, trial t^m)
// are isometric, |S_m y_b - x_a| = |y_b - S_m x_a|, so one evaluation of
// r, 1/r, sin and cos at (a, b) serves both blocks:
//   (t, s^m):  d = y_b - S_m x_a,  proj =  d . n_s
//   (s, t^m):                      proj = -d . (S_m n_t)
// (S_m is its own inverse and orthogonal).  The two 3x3 blocks are summed in
// two LDS images of the tiles' (rows x cols) sub-matrices with LDS float
// atomics and leave with one global atomic add per touched entry.  On a
// diagonal block (T == S) lane s does the pair {t, s} only for s >= t, the
// second direction only for s > t.  Images run in a runtime loop (mirror
// signs applied to the uniform test points).
// ---------------------------------------------------------------------------

constexpr int kTestsPerWave = kTileTris / 4;

// Both directions of one (t, s, image) combination over the 6 x 6 rule,
// added to dF[i][j] (test basis i on t, trial basis j on s) and dR[i][j]
// (test basis i on s, trial basis j on t) as -4 pi DLP / jac_s; sF[i]
// (basis on t) and sR[i] (basis on s) get 4 pi SLP / jac_s with the
// source's unit density.  tsc = 2 area(t) scales t's weights.
// kPrecise: the points are vertex 0 plus an offset (o = X e1 + Y e2, from
// exact float vertices), yp holds the trial offsets yo and xa_m the FP64
//   B_a = (v0_s - S_m v0_t) - S_m xo_a,   d = y_b - S_m x_a = yo_b + B_a,
// and the distance and the phase are formed in FP64 (every difference is
// then rounded relative to |d|, and k r keeps its accuracy at large k r).
// Both triangles are flat, so the normal projections of d = y_b - S_m x_a
// depend on one point only:
//   d . n_s         = (v0_s - S_m x_a) . n_s          = hF[a]
//   d . (S_m n_t)   = (y_b - S_m v0_t) . (S_m n_t)    = hR[b]
// (differences of nearby points first, so the rounding scales with |d|).
template <bool kComplexK, bool kSource, bool kPrecise, typename XT>
__device__ __forceinline__ void regular_both(const XT (&xa_m)[18], const float (&hF)[6],
                                             const float (&hR)[6], const float (&yp)[18],
                                             float k, float krev, double krev64, float kim, float tsc,
                                             float (&dF)[9][2],
                                             float (&dR)[9][2], float (&sF)[3][2],
                                             float (&sR)[3][2]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const XT xa = xa_m[3 * a], ya = xa_m[3 * a + 1], za = xa_m[3 * a + 2];
        float eF[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
        float eR[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
        float gF[2] = {0.f, 0.f};
        float gR[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
        for (int bq = 0; bq < 6; ++bq) {
            float r, rinv, s, c;
            if constexpr (kPrecise) {
                // d = yo_b + B_a in FP64; distance and phase in FP64
                const double dxd = (double)yp[3 * bq] + xa, dyd = (double)yp[3 * bq + 1] + ya,
                             dzd = (double)yp[3 * bq + 2] + za;
                const double r2d = fma(dxd, dxd, fma(dyd, dyd, dzd * dzd));
                const float r2 = fmaxf((float)r2d, 1.0e-14f);
                rinv = __builtin_amdgcn_rsqf(r2);
                r = r2 * rinv;
                const double rd = (double)r + fma(-(double)r, (double)r, r2d) * (double)(0.5f * rinv);
                double td = rd * krev64;
                td -= rint(td);
                s = __builtin_amdgcn_sinf((float)td);
                c = __builtin_amdgcn_cosf((float)td);
            } else {
                const float dx = yp[3 * bq] - xa;
                const float dy = yp[3 * bq + 1] - ya;
                const float dz = yp[3 * bq + 2] - za;
                const float r2 = fmaxf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), 1.0e-14f);
                rinv = __builtin_amdgcn_rsqf(r2);
                r = r2 * rinv;
                // phase in revolutions, reduced to [0, 1) (exact)
                const float t = __builtin_amdgcn_fractf(r * krev);
                s = __builtin_amdgcn_sinf(t);
                c = __builtin_amdgcn_cosf(t);
            }
            float q2 = rinv * rinv;  // 1/r^2 (times the damping)
            float damp = 1.0f;
            if (kComplexK) {
                damp = __expf(-kim * r);
                q2 *= damp;
            }
            const float fr = kComplexK ? rinv + kim : rinv;
            // -4 pi dG/dn = proj/r^2 * (c fr + s k, s fr - c k) e^{-kim r}
            const float p1 = fmaf(s, k, c * fr);
            const float p2 = fmaf(-c, k, s * fr);
            const float uF = hF[a] * q2;
            const float uR = -hR[bq] * q2;
            const float wb = (float)kRule6W[bq];
            const float w1 = wb * (float)kRule6X[bq];
            const float w2 = wb * (float)kRule6Y[bq];
            const float w0 = wb - w1 - w2;
            const float f1 = p1 * uF, f2 = p2 * uF;
            const float r1 = p1 * uR, r2v = p2 * uR;
            eF[0][0] = fmaf(f1, w0, eF[0][0]);
            eF[0][1] = fmaf(f2, w0, eF[0][1]);
            eF[1][0] = fmaf(f1, w1, eF[1][0]);
            eF[1][1] = fmaf(f2, w1, eF[1][1]);
            eF[2][0] = fmaf(f1, w2, eF[2][0]);
            eF[2][1] = fmaf(f2, w2, eF[2][1]);
            eR[0][0] = fmaf(r1, w0, eR[0][0]);
            eR[0][1] = fmaf(r2v, w0, eR[0][1]);
            eR[1][0] = fmaf(r1, w1, eR[1][0]);
            eR[1][1] = fmaf(r2v, w1, eR[1][1]);
            eR[2][0] = fmaf(r1, w2, eR[2][0]);
            eR[2][1] = fmaf(r2v, w2, eR[2][1]);
            if (kSource) {
                const float g = rinv * damp;
                const float gc = c * g, gs = s * g;
                gF[0] = fmaf(gc, wb, gF[0]);
                gF[1] = fmaf(gs, wb, gF[1]);
                gR[0][0] = fmaf(gc, w0, gR[0][0]);
                gR[0][1] = fmaf(gs, w0, gR[0][1]);
                gR[1][0] = fmaf(gc, w1, gR[1][0]);
                gR[1][1] = fmaf(gs, w1, gR[1][1]);
                gR[2][0] = fmaf(gc, w2, gR[2][0]);
                gR[2][1] = fmaf(gs, w2, gR[2][1]);
            }
        }
        const float wa = (float)kRule6W[a] * tsc;
        const float ta[3] = {wa * (float)(1.0 - kRule6X[a] - kRule6Y[a]), wa * (float)kRule6X[a],
                             wa * (float)kRule6Y[a]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                dF[3 * i + j][0] = fmaf(ta[i], eF[j][0], dF[3 * i + j][0]);
                dF[3 * i + j][1] = fmaf(ta[i], eF[j][1], dF[3 * i + j][1]);
                dR[3 * i + j][0] = fmaf(ta[j], eR[i][0], dR[3 * i + j][0]);
                dR[3 * i + j][1] = fmaf(ta[j], eR[i][1], dR[3 * i + j][1]);
            }
            if (kSource) {
                sF[i][0] = fmaf(ta[i], gF[0], sF[i][0]);
                sF[i][1] = fmaf(ta[i], gF[1], sF[i][1]);
                sR[i][0] = fmaf(wa, gR[i][0], sR[i][0]);
                sR[i][1] = fmaf(wa, gR[i][1], sR[i][1]);
            }
        }
    }
}

// The fields of a test triangle's records the regular kernel uses.  They are
// wave-uniform (scalar registers).  The next test's records are fetched one
// test ahead as one vector load spread over the lanes (lane l < 32: float
// word l, lane 32 + l: int word l) and moved to scalar registers with
// v_readlane, so waiting for them never waits for the LDS adds in flight
// (scalar loads share their counter with LDS operations).
struct TestRec {
    float x[18];   // quadrature points
    float n[3];    // unit normal
    float v0[3];   // vertex 0
    float e1[3];   // v1 - v0
    float e2[3];   // v2 - v0
    float jac;     // 2 * area
    int vid[3];    // vertex ids
    int key[3];    // coordinate-key ids under the identity
    int slot[3];   // tile dof slots
    int dof[3];    // P1 dofs
    int orig;      // caller index
    int dev;       // device index
};

static_assert(kTriF <= 32 && kTriI <= 32, "records must fit one wave");

__device__ __forceinline__ int rec_word(const float *__restrict__ tf,
                                        const int32_t *__restrict__ tii, int lane) {
    if (lane < kTriF) return __float_as_int(tf[lane]);
    if (lane >= 32 && lane < 32 + kTriI) return tii[lane - 32];
    return 0;
}

__device__ __forceinline__ void unpack_test(int w, TestRec &r) {
    auto f = [w](int j) { return __int_as_float(__builtin_amdgcn_readlane(w, j)); };
    auto i = [w](int j) { return __builtin_amdgcn_readlane(w, 32 + j); };
#pragma unroll
    for (int j = 0; j < 18; ++j) r.x[j] = f(j);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.n[c] = f(18 + c);
        r.v0[c] = f(23 + c);
        r.e1[c] = f(26 + c) - r.v0[c];
        r.e2[c] = f(29 + c) - r.v0[c];
        r.dof[c] = i(c);
        r.vid[c] = i(3 + c);
        r.key[c] = i(6 + c);
        r.slot[c] = i(20 + c);
    }
    r.jac = f(22);
    r.orig = i(19);
    r.dev = i(23);
}

// All images of one (t, s): tr is t's record (uniform), yp / v0s / ns the
// lane's points, vertex 0 and normal; rf (0 or 1) keeps or drops the (s, t)
// double-layer block.
template <bool kComplexK, bool kSource, int kM, bool kPrecise>
__device__ __forceinline__ void regular_images(const TestRec &tr, int n_masks,
                                               int packed_masks, unsigned skip,
                                               const float (&yp)[18], const float (&v0s)[3],
                                               const float (&ns)[3], float k, float krev,
                                               double krev64, float kim, float rf,
                                               float (&dF)[9][2],
                                               float (&dR)[9][2], float (&sF)[3][2],
                                               float (&sR)[3][2]) {
    // one or two images: unrolled; four: a runtime loop (code size)
    const int nm = kM < 4 ? kM : n_masks;
    for (int m = 0; m < nm; ++m) {
        if ((skip >> m) & 1u) continue;
        const int mask = (packed_masks >> (3 * m)) & 7;
        const float sg[3] = {(mask & 1) ? -1.0f : 1.0f, (mask & 2) ? -1.0f : 1.0f,
                             (mask & 4) ? -1.0f : 1.0f};
        float nt[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) nt[c] = sg[c] * tr.n[c];
        using XT = typename std::conditional<kPrecise, double, float>::type;
        XT xm[18];
        float hF[6], hR[6];
        if constexpr (kPrecise) {
            double av[3];
            float avf[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                av[c] = (double)v0s[c] - (double)(sg[c] * tr.v0[c]);  // exact
                avf[c] = (float)av[c];
            }
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                float h = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float xo = fmaf((float)kRule6X[a], tr.e1[c], (float)kRule6Y[a] * tr.e2[c]);
                    xm[3 * a + c] = av[c] - (double)(sg[c] * xo);
                    h = fmaf((float)xm[3 * a + c], ns[c], h);
                }
                hF[a] = h;
            }
#pragma unroll
            for (int bq = 0; bq < 6; ++bq) {
                float h = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) h = fmaf(yp[3 * bq + c] + avf[c], nt[c], h);
                hR[bq] = h * rf;  // rf = 0 drops the (s, t) double-layer block
            }
        } else {
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                float h = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    xm[3 * a + c] = sg[c] * tr.x[3 * a + c];
                    h = fmaf(v0s[c] - xm[3 * a + c], ns[c], h);
                }
                hF[a] = h;
            }
            float v0t[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) v0t[c] = sg[c] * tr.v0[c];
#pragma unroll
            for (int bq = 0; bq < 6; ++bq) {
                float h = 0.0f;
#pragma unroll
                for (int c = 0; c < 3; ++c) h = fmaf(yp[3 * bq + c] - v0t[c], nt[c], h);
                hR[bq] = h * rf;  // rf = 0 drops the (s, t) double-layer block
            }
        }
        regular_both<kComplexK, kSource, kPrecise>(xm, hF, hR, yp, k, krev, krev64, kim, tr.jac,
                                                   dF, dR, sF, sR);
    }
}

// FP32 path of one (t, s) over all images with the test points outer and
// the images inner: the image sums of a test point share one outer product
// with t's basis weights (dF / dR / sF / sR are linear in the image sums),
// so the 3 x 3 block accumulation costs 36 FMAs per test point instead of
// 36 per test point and image.  hR[m][b] are formed once per pair.  The
// kernel is issue-bound, so the per-evaluation work is kept to the terms
// that change with the trial point: q2 = 1/r^2 is folded into the phase
// terms (g = (sin k + cos / r) / r^2 ...), the test-side normal factor hF
// (fixed per test point and image) multiplies the image's sums once, and the
// RHS point sum is the sum of its three basis sums (the basis weights of a
// point add up to its weight).  Only non-touching combinations come here
// (touching ones are skipped or go to the singular kernel), so r > 0.
template <bool kComplexK, bool kSource, int kM>
__device__ __forceinline__ void regular_pair_fp32(const TestRec &tr, int n_masks, int packed_masks,
                                                  unsigned skip, const float (&yp)[18],
                                                  const float (&v0s)[3], const float (&ns)[3],
                                                  float k, float krev, float kim, float rf,
                                                  float (&dF)[9][2], float (&dR)[9][2],
                                                  float (&sF)[3][2], float (&sR)[3][2]) {
    const int nm = kM < 4 ? kM : n_masks;
    float sg[kM][3], hR[kM][6];
#pragma unroll
    for (int m = 0; m < kM; ++m) {
        const int mask = (packed_masks >> (3 * m)) & 7;
        sg[m][0] = (mask & 1) ? -1.0f : 1.0f;
        sg[m][1] = (mask & 2) ? -1.0f : 1.0f;
        sg[m][2] = (mask & 4) ? -1.0f : 1.0f;
        float nt[3], v0t[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            nt[c] = sg[m][c] * tr.n[c];
            v0t[c] = sg[m][c] * tr.v0[c];
        }
#pragma unroll
        for (int bq = 0; bq < 6; ++bq) {
            float h = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) h = fmaf(yp[3 * bq + c] - v0t[c], nt[c], h);
            hR[m][bq] = -h * rf;  // rf = 0 drops the (s, t) double-layer block
        }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        float eF[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
        float eR[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
        float gR[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
        for (int m = 0; m < kM; ++m) {
            if (m >= nm || ((skip >> m) & 1u)) continue;
            const float xa = sg[m][0] * tr.x[3 * a], ya = sg[m][1] * tr.x[3 * a + 1],
                        za = sg[m][2] * tr.x[3 * a + 2];
            const float hF = fmaf(v0s[0] - xa, ns[0], fmaf(v0s[1] - ya, ns[1], (v0s[2] - za) * ns[2]));
            float em[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};  // this image's sums before hF
#pragma unroll
            for (int bq = 0; bq < 6; ++bq) {
                const float dx = yp[3 * bq] - xa;
                const float dy = yp[3 * bq + 1] - ya;
                const float dz = yp[3 * bq + 2] - za;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const float t = __builtin_amdgcn_fractf(r * krev);  // revolutions, [0, 1)
                const float sn = __builtin_amdgcn_sinf(t);
                const float cs = __builtin_amdgcn_cosf(t);
                float q2 = rinv * rinv;
                float damp = 1.0f;
                if (kComplexK) {
                    damp = __expf(-kim * r);
                    q2 *= damp;
                }
                const float fr = kComplexK ? rinv + kim : rinv;
                // g = p q2 with p1 = sin k + cos fr, p2 = sin fr - cos k
                const float kq = k * q2, fq = fr * q2;
                const float g1 = fmaf(sn, kq, cs * fq);
                const float g2 = fmaf(-cs, kq, sn * fq);
                const float r1 = g1 * hR[m][bq], r2v = g2 * hR[m][bq];
                const float wb = (float)kRule6W[bq];
                const float w1 = wb * (float)kRule6X[bq];
                const float w2 = wb * (float)kRule6Y[bq];
                const float w0 = wb - w1 - w2;
                em[0][0] = fmaf(g1, w0, em[0][0]);
                em[0][1] = fmaf(g2, w0, em[0][1]);
                em[1][0] = fmaf(g1, w1, em[1][0]);
                em[1][1] = fmaf(g2, w1, em[1][1]);
                em[2][0] = fmaf(g1, w2, em[2][0]);
                em[2][1] = fmaf(g2, w2, em[2][1]);
                eR[0][0] = fmaf(r1, w0, eR[0][0]);
                eR[0][1] = fmaf(r2v, w0, eR[0][1]);
                eR[1][0] = fmaf(r1, w1, eR[1][0]);
                eR[1][1] = fmaf(r2v, w1, eR[1][1]);
                eR[2][0] = fmaf(r1, w2, eR[2][0]);
                eR[2][1] = fmaf(r2v, w2, eR[2][1]);
                if (kSource) {
                    const float gv = rinv * damp;
                    const float gc = cs * gv, gs = sn * gv;
                    gR[0][0] = fmaf(gc, w0, gR[0][0]);
                    gR[0][1] = fmaf(gs, w0, gR[0][1]);
                    gR[1][0] = fmaf(gc, w1, gR[1][0]);
                    gR[1][1] = fmaf(gs, w1, gR[1][1]);
                    gR[2][0] = fmaf(gc, w2, gR[2][0]);
                    gR[2][1] = fmaf(gs, w2, gR[2][1]);
                }
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                eF[j][0] = fmaf(hF, em[j][0], eF[j][0]);
                eF[j][1] = fmaf(hF, em[j][1], eF[j][1]);
            }
        }
        const float wa = (float)kRule6W[a] * tr.jac;
        const float ta[3] = {wa * (float)(1.0 - kRule6X[a] - kRule6Y[a]), wa * (float)kRule6X[a],
                             wa * (float)kRule6Y[a]};
        float gF[2] = {0.f, 0.f};
        if (kSource) {
            gF[0] = gR[0][0] + gR[1][0] + gR[2][0];
            gF[1] = gR[0][1] + gR[1][1] + gR[2][1];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                dF[3 * i + j][0] = fmaf(ta[i], eF[j][0], dF[3 * i + j][0]);
                dF[3 * i + j][1] = fmaf(ta[i], eF[j][1], dF[3 * i + j][1]);
                dR[3 * i + j][0] = fmaf(ta[j], eR[i][0], dR[3 * i + j][0]);
                dR[3 * i + j][1] = fmaf(ta[j], eR[i][1], dR[3 * i + j][1]);
            }
            if (kSource) {
                sF[i][0] = fmaf(ta[i], gF[0], sF[i][0]);
                sF[i][1] = fmaf(ta[i], gF[1], sF[i][1]);
                sR[i][0] = fmaf(wa, gR[i][0], sR[i][0]);
                sR[i][1] = fmaf(wa, gR[i][1], sR[i][1]);
            }
        }
    }
}

// One tile pair (pair index tp) at one frequency (b) of the regular
// assembly: the block body shared by the single-mesh kernel and the
// multi-mesh launch of a session group.
// list entries per (trial column, local vertex) of the gathered flush
constexpr int kGatherCap = 8;

template <bool kComplexK, int kM, bool kPrecise>
__device__ __forceinline__ void regular_tile_pair(
    const DeviceGeometry &g, int tp, int b, const float *__restrict__ k_real,
    const float *__restrict__ k_imag, const float *__restrict__ q, int n_src,
    const float *__restrict__ beta, float *__restrict__ A, int64_t lda, int64_t a_stride,
    float *__restrict__ rhs, int64_t rhs_stride) {
    // LDS tables, each [T dofs][S dofs] with real and imaginary planes:
    // F (rows of t, columns of s) and R (rows of s, columns of t), indexed
    // (test slot) * snd + (trial slot), so a wave's adds for one entry hit
    // consecutive words (one bank per distinct trial slot).
    extern __shared__ float acc[];
    const int T = g.n_triangles;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s_tile = g.tile_pairs[2 * tp], t_tile = g.tile_pairs[2 * tp + 1];
    const bool diag = s_tile == t_tile;
    const int s0 = g.tile_start[s_tile], scount = g.tile_count[s_tile], snd = g.tile_ndofs[s_tile];
    const int t0 = g.tile_start[t_tile], tcount = g.tile_count[t_tile], tnd = g.tile_ndofs[t_tile];
    const int plane = snd * tnd;
    float *accF = acc;              // re, then im at + plane
    float *accR = acc + 2 * plane;  // re, then im at + plane
    // single-image FP32 blocks gather each flush per trial column (below);
    // their tables hold the column lists first and are zeroed after
    constexpr bool kGather = kM == 1 && !kPrecise;
    if constexpr (!kGather)
        for (int e = threadIdx.x; e < 4 * snd * tnd; e += 256) acc[e] = 0.0f;

    const float k = k_real[b];
    const float kim = kComplexK ? k_imag[b] : 0.0f;
    const float krev = k * kInv2Pi;
    const double krev64 = (double)k * kInv2PiD;
    constexpr int M = kM;  // 1 + image masks (the session's n_masks)
    const int packed_masks = g.masks[0] | (g.masks[1] << 3) | (g.masks[2] << 6) | (g.masks[3] << 9);

    const bool trial_ok = lane < scount;
    const int tr = s0 + (trial_ok ? lane : 0);
    // the lane's points: absolute, or (kPrecise) offsets from vertex 0
    float yp[18], ns[3], v0s[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ns[c] = g.tri_f_soa[(size_t)(18 + c) * T + tr];
        v0s[c] = g.tri_f_soa[(size_t)(23 + c) * T + tr];
    }
    if constexpr (kPrecise) {
        float e1[3], e2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            e1[c] = g.tri_f_soa[(size_t)(26 + c) * T + tr] - v0s[c];
            e2[c] = g.tri_f_soa[(size_t)(29 + c) * T + tr] - v0s[c];
        }
#pragma unroll
        for (int b = 0; b < 6; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                yp[3 * b + c] = fmaf((float)kRule6X[b], e1[c], (float)kRule6Y[b] * e2[c]);
    } else {
#pragma unroll
        for (int j = 0; j < 18; ++j) yp[j] = g.tri_f_soa[(size_t)j * T + tr];
    }
    const float jac_s = g.tri_f_soa[(size_t)22 * T + tr];
    int svid[3], skid[kMaxMasks][3], slc[3], sdof[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        svid[l] = g.tri_i_soa[(size_t)(3 + l) * T + tr];
        slc[l] = g.tri_i_soa[(size_t)(20 + l) * T + tr];
        sdof[l] = g.tri_i_soa[(size_t)l * T + tr];
    }
#pragma unroll
    for (int m = 0; m < kMaxMasks; ++m)
#pragma unroll
        for (int l = 0; l < 3; ++l)
            skid[m][l] = (m < M) ? g.tri_i_soa[(size_t)(6 + 3 * m + l) * T + tr] : -1;
    (void)skid;
    const int orig = g.tri_i_soa[(size_t)19 * T + tr];
    // Neumann data of the lane's triangle for every source (q [B][S][T])
    bool lane_q = false;
    for (int src = 0; src < n_src; ++src) {
        const size_t qi = (((size_t)b * n_src + src) * T + orig) * 2;
        lane_q |= trial_ok && (q[qi] != 0.0f || q[qi + 1] != 0.0f);
    }
    // Robin admittance of the lane's triangle: its single layer enters the
    // matrix as -(i k beta / 3) SLP_i in every column of the triangle
    // (main.swift:4101-4110); i k = (-k_im, k_re)
    const float bre = (beta && trial_ok) ? beta[((size_t)b * T + orig) * 2] : 0.0f;
    const float bim = (beta && trial_ok) ? beta[((size_t)b * T + orig) * 2 + 1] : 0.0f;
    const float cFre = (-kim * bre - k * bim) * (1.0f / 3.0f);
    const float cFim = (-kim * bim + k * bre) * (1.0f / 3.0f);
    const bool lane_robin = bre != 0.0f || bim != 0.0f;
    const bool wave_has_source = __any(lane_q || lane_robin);
    // -4 pi and 4 pi folded out of the kernels, jac = (2 A_test)(2 A_trial)
    const float dscale = -kInv4Pi * jac_s * g.row_weight;
    const float gscale = kInv4Pi * jac_s * g.row_weight;
    // Column lists of the trial side (kGather): for trial dof slot c and
    // local vertex j, the lanes whose vertex j has slot c (at most kGatherCap;
    // more and the block keeps per-lane atomics).  A lane owns columns
    // lane and lane + 64 and holds their lists in registers, packed one byte
    // per entry.
    uint32_t glist[2][3][2] = {};
    int gcnt[2][3] = {};
    int gmax[3] = {0, 0, 0};
    bool gather = false;
    if constexpr (kGather) {
        int *cnt = reinterpret_cast<int *>(acc);                            // [3][snd]
        unsigned char *lst = reinterpret_cast<unsigned char *>(cnt + 3 * snd);  // [3][snd][cap]
        for (int e = threadIdx.x; e < 3 * snd; e += 256) cnt[e] = 0;
        __syncthreads();
        if (wave == 0 && trial_ok) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int slot = j * snd + slc[j];
                const int pos = atomicAdd(cnt + slot, 1);
                if (pos < kGatherCap) lst[slot * kGatherCap + pos] = (unsigned char)lane;
            }
        }
        __syncthreads();
        int over = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = lane + 64 * h;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int n = c < snd ? cnt[j * snd + c] : 0;
                over |= n > kGatherCap;
                gcnt[h][j] = min(n, kGatherCap);
                if (n > 0) {
                    const uint32_t *w = reinterpret_cast<const uint32_t *>(lst) + (j * snd + c) * 2;
                    glist[h][j][0] = w[0];
                    glist[h][j][1] = w[1];
                }
            }
        }
        // every lane has read the lists; the flag is block-uniform
        gather = __syncthreads_or(over) == 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int m = max(gcnt[0][j], gcnt[1][j]);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off));
            gmax[j] = __builtin_amdgcn_readfirstlane(m);
        }
        for (int e = threadIdx.x; e < 4 * snd * tnd; e += 256) acc[e] = 0.0f;
    }
    (void)glist;
    (void)gcnt;
    (void)gmax;
    (void)gather;
    __syncthreads();

    const int n_tests = max(0, min(kTestsPerWave, tcount - wave * kTestsPerWave));
    // Per-test source flags and Robin admittances of this wave's tests, one
    // lane per test, read back with v_readlane inside the loop (no memory
    // access on the per-test path).
    bool tq_lane = false;
    float tbre_lane = 0.0f, tbim_lane = 0.0f;
    const int rbase = t0 + wave * kTestsPerWave;  // this wave's first test-side record
    if (lane < n_tests) {
        const int torig_l = g.tri_t_i[(size_t)(rbase + lane) * kTriI + 19];
        for (int src = 0; src < n_src; ++src) {
            const size_t qi = (((size_t)b * n_src + src) * T + torig_l) * 2;
            tq_lane |= q[qi] != 0.0f || q[qi + 1] != 0.0f;
        }
        if (beta) {
            tbre_lane = beta[((size_t)b * T + torig_l) * 2];
            tbim_lane = beta[((size_t)b * T + torig_l) * 2 + 1];
        }
    }
    const unsigned long long tq_mask = __ballot(tq_lane);
    // dF[3 r + j] / dR[3 i + r]: sums for the dof at local position r of the
    // current test, kept across tests while that position's dof is
    // unchanged (the test-side records keep shared dofs in place), and added
    // to the LDS tables only when it changes and after the last test.
    float dF[9][2], dR[9][2];
#pragma unroll
    for (int i = 0; i < 9; ++i) dF[i][0] = dF[i][1] = dR[i][0] = dR[i][1] = 0.0f;
    int held[3] = {-1, -1, -1};  // tile dof slot of each position
    auto flush = [&](int r) {
        if (kGather && gather) {
            // lanes own trial columns: each sums the contributions of the
            // lanes in its lists (cross-lane reads), then one atomic per
            // table word and owned column
            const int row = held[r];
            float v[3][4];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                v[j][0] = trial_ok ? dF[3 * r + j][0] * dscale : 0.0f;
                v[j][1] = trial_ok ? dF[3 * r + j][1] * dscale : 0.0f;
                v[j][2] = trial_ok ? dR[3 * j + r][0] * dscale : 0.0f;
                v[j][3] = trial_ok ? dR[3 * j + r][1] * dscale : 0.0f;
            }
            const int nh = snd > 64 ? 2 : 1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (h >= nh) break;
                float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < gmax[j]; ++k) {
                        const bool on = k < gcnt[h][j];
                        const int src =
                            on ? (int)((glist[h][j][k >> 2] >> (8 * (k & 3))) & 0xffu) : lane;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const float x = __shfl(v[j][q], src);
                            sum[q] += on ? x : 0.0f;
                        }
                    }
                const int c = lane + 64 * h;
                if (c < snd && (gcnt[h][0] + gcnt[h][1] + gcnt[h][2]) > 0) {
                    float *cF = accF + row * snd + c;
                    float *cR = accR + row * snd + c;
                    __hip_atomic_fetch_add(cF, sum[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(cF + plane, sum[1], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(cR, sum[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(cR + plane, sum[3], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        } else if (trial_ok) {
            float *cF = accF + held[r] * snd;
            float *cR = accR + held[r] * snd;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // LDS table adds: relaxed, workgroup scope (no ordering among them)
                __hip_atomic_fetch_add(cF + slc[j], dF[3 * r + j][0] * dscale, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(cF + slc[j] + plane, dF[3 * r + j][1] * dscale,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(cR + slc[j], dR[3 * j + r][0] * dscale, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(cR + slc[j] + plane, dR[3 * j + r][1] * dscale,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j)
            dF[3 * r + j][0] = dF[3 * r + j][1] = dR[3 * j + r][0] = dR[3 * j + r][1] = 0.0f;
    };
    int rec = 0;
    if (n_tests > 0)
        rec = rec_word(g.tri_t_f + (size_t)rbase * kTriF, g.tri_t_i + (size_t)rbase * kTriI, lane);
    for (int ti = 0; ti < n_tests; ++ti) {
        TestRec cur;
        unpack_test(rec, cur);
        // next test's record in flight while this one is evaluated
        const int nt_next = rbase + ti + (ti + 1 < n_tests ? 1 : 0);
        rec = rec_word(g.tri_t_f + (size_t)nt_next * kTriF, g.tri_t_i + (size_t)nt_next * kTriI,
                       lane);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            if (held[r] != cur.slot[r]) {
                if (held[r] >= 0) flush(r);
                held[r] = cur.slot[r];
            }
        const int lt = cur.dev - t0;  // the test's position in its tile
        const int torig = cur.orig;
        const bool test_q = (tq_mask >> ti) & 1ull;
        const float tbre = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tbre_lane), ti));
        const float tbim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tbim_lane), ti));
        const bool test_robin = tbre != 0.0f || tbim != 0.0f;
        const bool test_source = test_q || test_robin;
        // lanes and directions of this (t, s) pair
        const bool fwd = trial_ok && (!diag || lane >= lt);
        const bool rev = trial_ok && (!diag || lane > lt);
        // skip the singular combinations: t touches s^m (image 0 by vertex
        // index, images by coordinate key of t against the keys of s under
        // mask m); s touches t^m exactly when t touches s^m
        unsigned skip = fwd ? 0u : 0xfu;
#pragma unroll
        for (int m = 0; m < kM; ++m) {
            bool touch = false;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    touch |= (m == 0) ? (cur.vid[i] == svid[j]) : (cur.key[i] == skid[m][j]);
            if (m < M && touch) skip |= 1u << m;
        }
        float sF[3][2], sR[3][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) sF[i][0] = sF[i][1] = sR[i][0] = sR[i][1] = 0.0f;
        const float rf = rev ? 1.0f : 0.0f;
        const bool src = wave_has_source || test_source;
        if constexpr (kPrecise) {
            if (src)
                regular_images<kComplexK, true, kM, true>(cur, g.n_masks, packed_masks, skip, yp,
                                                          v0s, ns, k, krev, krev64, kim, rf, dF,
                                                          dR, sF, sR);
            else
                regular_images<kComplexK, false, kM, true>(cur, g.n_masks, packed_masks, skip, yp,
                                                           v0s, ns, k, krev, krev64, kim, rf, dF,
                                                           dR, sF, sR);
        } else {
            if (src)
                regular_pair_fp32<kComplexK, true, kM>(cur, g.n_masks, packed_masks, skip, yp, v0s,
                                                       ns, k, krev, kim, rf, dF, dR, sF, sR);
            else
                regular_pair_fp32<kComplexK, false, kM>(cur, g.n_masks, packed_masks, skip, yp,
                                                        v0s, ns, k, krev, kim, rf, dF, dR, sF, sR);
        }
        if (src) {
            // Robin fold in the units of dF / dR (scaled by dscale when they
            // reach the tables): -(i k beta / 3) SLP = dscale * (+(i k beta / 3) s)
            if (lane_robin) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float are = cFre * sF[i][0] - cFim * sF[i][1];
                    const float aim = cFre * sF[i][1] + cFim * sF[i][0];
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        dF[3 * i + j][0] += are;
                        dF[3 * i + j][1] += aim;
                    }
                }
            }
            if (test_robin && rev) {
                const float cRre = (-kim * tbre - k * tbim) * (1.0f / 3.0f);
                const float cRim = (-kim * tbim + k * tbre) * (1.0f / 3.0f);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float are = cRre * sR[i][0] - cRim * sR[i][1];
                    const float aim = cRre * sR[i][1] + cRim * sR[i][0];
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        dR[3 * i + j][0] += are;
                        dR[3 * i + j][1] += aim;
                    }
                }
            }
        }
        if (src) {
            const float gsc = gscale;  // sF / sR carry 2 area(t)
            for (int sv = 0; sv < n_src; ++sv) {
                float *rb = rhs + ((size_t)b * n_src + sv) * rhs_stride * 2;
                const size_t lq = (((size_t)b * n_src + sv) * T + orig) * 2;
                const float qre = trial_ok ? q[lq] : 0.0f, qim = trial_ok ? q[lq + 1] : 0.0f;
                if (__any(qre != 0.0f || qim != 0.0f)) {
                    // rhs_t += 2^planes SLP(t, s) q_s, reduced over the wave
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        float cre = fwd ? (sF[i][0] * qre - sF[i][1] * qim) * gsc : 0.0f;
                        float cim = fwd ? (sF[i][0] * qim + sF[i][1] * qre) * gsc : 0.0f;
#pragma unroll
                        for (int off = 32; off >= 1; off >>= 1) {
                            cre += __shfl_xor(cre, off);
                            cim += __shfl_xor(cim, off);
                        }
                        if (lane == 0) {
                            atomicAdd(rb + (size_t)cur.dof[i] * 2, cre);
                            atomicAdd(rb + (size_t)cur.dof[i] * 2 + 1, cim);
                        }
                    }
                }
                const size_t tq = (((size_t)b * n_src + sv) * T + torig) * 2;
                const float tqre = q[tq], tqim = q[tq + 1];
                if ((tqre != 0.0f || tqim != 0.0f) && rev) {
                    // rhs_s += 2^planes SLP(s, t) q_t
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        atomicAdd(rb + (size_t)sdof[i] * 2, (sR[i][0] * tqre - sR[i][1] * tqim) * gsc);
                        atomicAdd(rb + (size_t)sdof[i] * 2 + 1, (sR[i][0] * tqim + sR[i][1] * tqre) * gsc);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        if (held[r] >= 0) flush(r);
    __syncthreads();
    float *Ab = A + (size_t)b * a_stride * 2;
    const int32_t *tdofs = g.tile_dofs + (size_t)t_tile * kTileMaxDofs;
    const int32_t *sdofs = g.tile_dofs + (size_t)s_tile * kTileMaxDofs;
    // consecutive threads take consecutive rows of one matrix column (the
    // atomics of a wave share cache lines)
    for (int e = threadIdx.x; e < plane; e += 256) {
        const int ls = e / tnd, lt_ = e - ls * tnd;
        const float fre = accF[lt_ * snd + ls], fim = accF[lt_ * snd + ls + plane];
        if (fre != 0.0f || fim != 0.0f) {  // A[t dof][s dof]
            const size_t idx = ((size_t)sdofs[ls] * lda + tdofs[lt_]) * 2;
            atomicAdd(Ab + idx, fre);
            atomicAdd(Ab + idx + 1, fim);
        }
    }
    for (int e = threadIdx.x; e < plane; e += 256) {
        const int lt_ = e / snd, ls = e - lt_ * snd;
        const float rre = accR[e], rim = accR[e + plane];
        if (rre != 0.0f || rim != 0.0f) {  // A[s dof][t dof]
            const size_t idx = ((size_t)tdofs[lt_] * lda + sdofs[ls]) * 2;
            atomicAdd(Ab + idx, rre);
            atomicAdd(Ab + idx + 1, rim);
        }
    }
}

template <bool kComplexK, int kM, bool kPrecise>
__global__ __launch_bounds__(256, 2) void k_assemble_regular(
    DeviceGeometry g, const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ q, int n_src, const float *__restrict__ beta,
    float *__restrict__ A, int64_t lda, int64_t a_stride, float *__restrict__ rhs,
    int64_t rhs_stride) {
    regular_tile_pair<kComplexK, kM, kPrecise>(g, blockIdx.x, blockIdx.y, k_real, k_imag, q, n_src,
                                               beta, A, lda, a_stride, rhs, rhs_stride);
}

// The member of a multi-mesh launch that owns flattened block `blk`: jobs
// are ordered by their first block (a handful of members: a scalar scan).
__device__ __forceinline__ int job_of_block(const AssemblyJob *__restrict__ jobs, int n_jobs,
                                            int blk, bool singular) {
    int j = 0;
    for (int i = 1; i < n_jobs; ++i)
        if ((singular ? jobs[i].block_begin_sing : jobs[i].block_begin_reg) <= blk) j = i;
    return j;
}

// Regular assembly of every member of a session group in one launch: block
// x = (member, frequency, tile pair) flattened; members share the mask count
// kM and the real / complex wavenumber choice.
template <bool kComplexK, int kM>
__global__ __launch_bounds__(256, 2) void k_assemble_regular_multi(
    const AssemblyJob *__restrict__ jobs, int n_jobs) {
    const int j = job_of_block(jobs, n_jobs, blockIdx.x, false);
    const AssemblyJob &jb = jobs[j];
    const int local = blockIdx.x - jb.block_begin_reg;
    const int ntp = jb.g.n_tile_pairs;
    regular_tile_pair<kComplexK, kM, false>(jb.g, local % ntp, local / ntp, jb.k_real, jb.k_imag,
                                            jb.q, jb.n_src, jb.beta, jb.A, jb.lda, jb.a_stride,
                                            jb.rhs, jb.rhs_stride);
}

// ---------------------------------------------------------------------------
// Singular pair classes.  One lane per class, kFB frequencies per lane; a
// wave holds classes of one kind, so the Sauter-Schwab points are
// wave-uniform and stream through scalar loads.  Per point and lane:
// the distance vector from the class's affine maps (cancellation-free on
// shared vertices and edges), one rsq, then per frequency sin/cos and the
// fused 3x3 double-layer block plus 3 single-layer moments.  Sums run in
// FP32 over runs of kSgRun points and are carried in FP64 across runs
// (blocked compensated summation).  Block deltas go to A / rhs with one
// atomic add per entry.
// ---------------------------------------------------------------------------

constexpr int kSgRun = 32;

// Block (bx, by) of the singular-class assembly (4 waves of classes, kFB
// frequencies): the body shared by the single-mesh kernel and the
// multi-mesh launch.
template <bool kComplexK, int kFB>
__device__ __forceinline__ void singular_block(
    const DeviceGeometry &g, int bx, int by, const float *__restrict__ rule, int n_batch,
    const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ q, int n_src, const float *__restrict__ beta,
    float *__restrict__ A, int64_t lda, int64_t a_stride, float *__restrict__ rhs,
    int64_t rhs_stride) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(bx * 4 + (threadIdx.x >> 6));
    const int w1 = g.sg_waves[0], w2 = w1 + g.sg_waves[1], w3 = w2 + g.sg_waves[2];
    if (w >= w3) return;
    const int kind = w < w1 ? 0 : (w < w2 ? 1 : 2);
    const int np = g.sg_np;
    const int c = w * 64 + lane;
    const float *__restrict__ F = g.sg_f;
    float gf[18];
#pragma unroll
    for (int j = 0; j < 18; ++j) gf[j] = F[(size_t)j * np + c];
    const float r0x = gf[0], r0y = gf[1], r0z = gf[2];
    const float e1x = gf[3], e1y = gf[4], e1z = gf[5];
    const float e2x = gf[6], e2y = gf[7], e2z = gf[8];
    const float g1x = gf[9], g1y = gf[10], g1z = gf[11];
    const float g2x = gf[12], g2y = gf[13], g2z = gf[14];
    const float nx = gf[15], ny = gf[16], nz = gf[17];

    const int b0 = by * kFB;
    float kk[kFB], krev[kFB], kim[kFB];
#pragma unroll
    for (int f = 0; f < kFB; ++f) {
        const int b = min(b0 + f, n_batch - 1);
        kk[f] = k_real[b];
        krev[f] = kk[f] * kInv2Pi;
        kim[f] = kComplexK ? k_imag[b] : 0.0f;
    }

    const float *__restrict__ rp = rule + (size_t)g.sg_rule_offset[kind] * kSgRule;
    const int cnt = g.sg_rule_count[kind];
    double dd[kFB][9][2], ds[kFB][3][2];
#pragma unroll
    for (int f = 0; f < kFB; ++f) {
#pragma unroll
        for (int i = 0; i < 9; ++i) dd[f][i][0] = dd[f][i][1] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) ds[f][i][0] = ds[f][i][1] = 0.0;
    }
    for (int p0 = 0; p0 < cnt; p0 += kSgRun) {
        float ad[kFB][9][2], as[kFB][3][2];
#pragma unroll
        for (int f = 0; f < kFB; ++f) {
#pragma unroll
            for (int i = 0; i < 9; ++i) ad[f][i][0] = ad[f][i][1] = 0.0f;
#pragma unroll
            for (int i = 0; i < 3; ++i) as[f][i][0] = as[f][i][1] = 0.0f;
        }
        // Two-stage software pipeline over the points: the scalar loads of
        // the next point are in flight while the current one is evaluated.
        // distance vector of a point (its first use of the point's loads)
        auto dist = [&](const float (&pt)[kSgRule], float &rx, float &ry, float &rz) {
            const float dx = pt[0], dy = pt[1], xs = pt[2], ys = pt[3];
            rx = fmaf(ys, g2x, fmaf(xs, g1x, fmaf(dy, e2x, fmaf(dx, e1x, r0x))));
            ry = fmaf(ys, g2y, fmaf(xs, g1y, fmaf(dy, e2y, fmaf(dx, e1y, r0y))));
            rz = fmaf(ys, g2z, fmaf(xs, g1z, fmaf(dy, e2z, fmaf(dx, e1z, r0z))));
        };
        auto eval_point = [&](const float (&pt)[kSgRule], float rx, float ry, float rz) {
            const float r2raw = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
            const bool ok = r2raw > 1.0e-14f;
            const float r2 = ok ? r2raw : 1.0e-14f;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            const float gs = ok ? kInv4Pi * rinv : 0.0f;
            const float proj = fmaf(rx, nx, fmaf(ry, ny, rz * nz)) * rinv;
#pragma unroll
            for (int f = 0; f < kFB; ++f) {
                float t = r * krev[f];
                t = t - __builtin_rintf(t);
                const float sn = __builtin_amdgcn_sinf(t);
                const float cs = __builtin_amdgcn_cosf(t);
                float sc = gs;
                if (kComplexK) sc *= __expf(-kim[f] * r);
                const float gre = cs * sc, gim = sn * sc;
                const float fre = kComplexK ? (-rinv - kim[f]) : -rinv;
                const float pre = gre * proj, pim = gim * proj;
                const float dre = fmaf(pre, fre, -pim * kk[f]);
                const float dim = fmaf(pre, kk[f], pim * fre);
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    ad[f][i][0] = fmaf(dre, pt[4 + i], ad[f][i][0]);
                    ad[f][i][1] = fmaf(dim, pt[4 + i], ad[f][i][1]);
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    as[f][i][0] = fmaf(gre, pt[13 + i], as[f][i][0]);
                    as[f][i][1] = fmaf(gim, pt[13 + i], as[f][i][1]);
                }
            }
        };
        float pa[kSgRule], pb[kSgRule];
#pragma unroll
        for (int j = 0; j < kSgRule; ++j) pa[j] = rp[(size_t)p0 * kSgRule + j];
        // Scalar loads return out of order, so a wait on one point's loads
        // is a wait on all: each point's loads are consumed (dist) before
        // the next point's loads are issued, and the scheduling barriers keep
        // that order.
        for (int p = p0; p < p0 + kSgRun; p += 2) {
            float rx, ry, rz;
            dist(pa, rx, ry, rz);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < kSgRule; ++j) pb[j] = rp[(size_t)(p + 1) * kSgRule + j];
            __builtin_amdgcn_sched_barrier(0);
            eval_point(pa, rx, ry, rz);
            __builtin_amdgcn_sched_barrier(0);
            dist(pb, rx, ry, rz);
            __builtin_amdgcn_sched_barrier(0);
            const int pn = (p + 2 < cnt) ? p + 2 : p;
#pragma unroll
            for (int j = 0; j < kSgRule; ++j) pa[j] = rp[(size_t)pn * kSgRule + j];
            __builtin_amdgcn_sched_barrier(0);
            eval_point(pb, rx, ry, rz);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int f = 0; f < kFB; ++f) {
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                dd[f][i][0] += (double)ad[f][i][0];
                dd[f][i][1] += (double)ad[f][i][1];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                ds[f][i][0] += (double)as[f][i][0];
                ds[f][i][1] += (double)as[f][i][1];
            }
        }
    }

    const int orig = g.sg_i[(size_t)6 * np + c];
    if (orig < 0) return;
    const int T = g.n_triangles;
    const double scale = (double)F[(size_t)18 * np + c];
    const float mc = F[(size_t)19 * np + c];
    int rows[3], cols[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rows[a] = g.sg_i[(size_t)a * np + c];
        cols[a] = g.sg_i[(size_t)(3 + a) * np + c];
    }
#pragma unroll
    for (int f = 0; f < kFB; ++f) {
        const int b = b0 + f;
        if (b >= n_batch) break;
        float *Ab = A + (size_t)b * a_stride * 2;
        // Robin fold -(i k beta / 3) SLP_a of the trial triangle
        double cre = 0.0, cim = 0.0;
        if (beta) {
            const double br = beta[((size_t)b * T + orig) * 2];
            const double bi = beta[((size_t)b * T + orig) * 2 + 1];
            cre = (-(double)kim[f] * br - (double)kk[f] * bi) / 3.0;
            cim = (-(double)kim[f] * bi + (double)kk[f] * br) / 3.0;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
                const double rre = cre * ds[f][a][0] - cim * ds[f][a][1];
                const double rim = cre * ds[f][a][1] + cim * ds[f][a][0];
                float vre = (float)((dd[f][3 * a + bb][0] - rre) * scale);
                const float vim = (float)((dd[f][3 * a + bb][1] - rim) * scale);
                if (mc != 0.0f) vre += mc * (rows[a] == cols[bb] ? (1.0f / 6.0f) : (1.0f / 12.0f));
                const size_t idx = ((size_t)cols[bb] * lda + rows[a]) * 2;
                atomicAdd(Ab + idx, vre);
                atomicAdd(Ab + idx + 1, vim);
            }
        for (int sv = 0; sv < n_src; ++sv) {
            const size_t qi = (((size_t)b * n_src + sv) * T + orig) * 2;
            const float qre = q[qi], qim = q[qi + 1];
            if (qre == 0.0f && qim == 0.0f) continue;
            float *rb = rhs + ((size_t)b * n_src + sv) * rhs_stride * 2;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float sre = (float)(ds[f][a][0] * scale);
                const float sim = (float)(ds[f][a][1] * scale);
                atomicAdd(rb + (size_t)rows[a] * 2, sre * qre - sim * qim);
                atomicAdd(rb + (size_t)rows[a] * 2 + 1, sre * qim + sim * qre);
            }
        }
    }
}

template <bool kComplexK, int kFB>
__global__ __launch_bounds__(256) void k_assemble_singular(
    DeviceGeometry g, const float *__restrict__ rule, int n_batch,
    const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ q, int n_src, const float *__restrict__ beta,
    float *__restrict__ A, int64_t lda, int64_t a_stride, float *__restrict__ rhs,
    int64_t rhs_stride) {
    singular_block<kComplexK, kFB>(g, blockIdx.x, blockIdx.y, rule, n_batch, k_real, k_imag, q,
                                   n_src, beta, A, lda, a_stride, rhs, rhs_stride);
}

// Singular classes of every member of a session group in one launch: block
// x = (member, frequency group, wave block) flattened.
template <bool kComplexK, int kFB>
__global__ __launch_bounds__(256) void k_assemble_singular_multi(
    const AssemblyJob *__restrict__ jobs, int n_jobs) {
    const int j = job_of_block(jobs, n_jobs, blockIdx.x, true);
    const AssemblyJob &jb = jobs[j];
    const int local = blockIdx.x - jb.block_begin_sing;
    const int waves = jb.g.sg_waves[0] + jb.g.sg_waves[1] + jb.g.sg_waves[2];
    const int nbx = (waves + 3) / 4;
    singular_block<kComplexK, kFB>(jb.g, local % nbx, local / nbx, jb.g.sg_rule, jb.n_batch,
                                   jb.k_real, jb.k_imag, jb.q, jb.n_src, jb.beta, jb.A, jb.lda,
                                   jb.a_stride, jb.rhs, jb.rhs_stride);
}

// ---------------------------------------------------------------------------
// Near-field corrections (applyNearFieldCorrectionsCPU,
// main.swift:3072-3175, with subdividedPairBlocks, main.swift:2840-2910):
// for every near class, delta = (level-L subdivided rule) - (6 x 6 rule) of
// the (test, trial image) block, added to A (and the Robin fold) and to the
// right-hand sides.  One lane per class, grid.y = frequency.  Both rules are
// invariant under the swap of the reference coordinates, so the image
// parity of the test side drops out and test points use the identity map.
// Per test point the trial sums are FP32; across test points FP64.
// ---------------------------------------------------------------------------
template <bool kComplexK>
__global__ __launch_bounds__(256) void k_assemble_near(
    DeviceGeometry g, const float *__restrict__ k_real, const float *__restrict__ k_imag,
    const float *__restrict__ q, int n_src, const float *__restrict__ beta,
    float *__restrict__ A, int64_t lda, int64_t a_stride, float *__restrict__ rhs,
    int64_t rhs_stride) {
    const int np = g.nc_np;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= np) return;
    const int orig = g.nc_i[(size_t)6 * np + c];
    if (orig < 0) return;
    const float *__restrict__ F = g.nc_f;
    float gf[18];
#pragma unroll
    for (int j = 0; j < 18; ++j) gf[j] = F[(size_t)j * np + c];
    const int b = blockIdx.y;
    const float kk = k_real[b];
    const float krev = kk * kInv2Pi;
    const float kim = kComplexK ? k_imag[b] : 0.0f;
    const float *__restrict__ rule = g.nc_rule;

    // sums[0]: subdivided rule, sums[1]: plain rule
    double ds[2][3][2], dd[2][9][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int i = 0; i < 3; ++i) ds[h][i][0] = ds[h][i][1] = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) dd[h][i][0] = dd[h][i][1] = 0.0;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float *__restrict__ rr = h == 0 ? rule : rule + (size_t)g.nc_sub_points * 3;
        const int n_pts = h == 0 ? g.nc_sub_points : 6;
        for (int a = 0; a < n_pts; ++a) {
            const float xa = rr[3 * a], ya = rr[3 * a + 1], wa = rr[3 * a + 2];
            // R0 - x E1 - y E2
            const float bx = fmaf(-ya, gf[6], fmaf(-xa, gf[3], gf[0]));
            const float by = fmaf(-ya, gf[7], fmaf(-xa, gf[4], gf[1]));
            const float bz = fmaf(-ya, gf[8], fmaf(-xa, gf[5], gf[2]));
            float sg[2] = {0.0f, 0.0f}, sd[3][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
            for (int bb = 0; bb < n_pts; ++bb) {
                const float xs = rr[3 * bb], ys = rr[3 * bb + 1], wb = rr[3 * bb + 2];
                const float dx = fmaf(ys, gf[12], fmaf(xs, gf[9], bx));
                const float dy = fmaf(ys, gf[13], fmaf(xs, gf[10], by));
                const float dz = fmaf(ys, gf[14], fmaf(xs, gf[11], bz));
                const GD v = kernel_gd<kComplexK>(dx, dy, dz, gf[15], gf[16], gf[17], kk, krev, kim);
                sg[0] = fmaf(wb, v.gre, sg[0]);
                sg[1] = fmaf(wb, v.gim, sg[1]);
                const float s0 = wb * (1.0f - xs - ys), s1 = wb * xs, s2 = wb * ys;
                sd[0][0] = fmaf(s0, v.dre, sd[0][0]);
                sd[0][1] = fmaf(s0, v.dim, sd[0][1]);
                sd[1][0] = fmaf(s1, v.dre, sd[1][0]);
                sd[1][1] = fmaf(s1, v.dim, sd[1][1]);
                sd[2][0] = fmaf(s2, v.dre, sd[2][0]);
                sd[2][1] = fmaf(s2, v.dim, sd[2][1]);
            }
            const double t[3] = {(double)(wa * (1.0f - xa - ya)), (double)(wa * xa), (double)(wa * ya)};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                ds[h][i][0] += t[i] * (double)sg[0];
                ds[h][i][1] += t[i] * (double)sg[1];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    dd[h][3 * i + j][0] += t[i] * (double)sd[j][0];
                    dd[h][3 * i + j][1] += t[i] * (double)sd[j][1];
                }
            }
        }
    }

    const int T = g.n_triangles;
    const double scale = (double)F[(size_t)18 * np + c];
    int rows[3], cols[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rows[a] = g.nc_i[(size_t)a * np + c];
        cols[a] = g.nc_i[(size_t)(3 + a) * np + c];
    }
    double sl[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        sl[i][0] = (ds[0][i][0] - ds[1][i][0]) * scale;
        sl[i][1] = (ds[0][i][1] - ds[1][i][1]) * scale;
    }
    float *Ab = A + (size_t)b * a_stride * 2;
    double cre = 0.0, cim = 0.0;
    if (beta) {
        const double br = beta[((size_t)b * T + orig) * 2];
        const double bi = beta[((size_t)b * T + orig) * 2 + 1];
        cre = (-(double)kim * br - (double)kk * bi) / 3.0;
        cim = (-(double)kim * bi + (double)kk * br) / 3.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double vre = (dd[0][3 * i + j][0] - dd[1][3 * i + j][0]) * scale -
                               (cre * sl[i][0] - cim * sl[i][1]);
            const double vim = (dd[0][3 * i + j][1] - dd[1][3 * i + j][1]) * scale -
                               (cre * sl[i][1] + cim * sl[i][0]);
            const size_t idx = ((size_t)cols[j] * lda + rows[i]) * 2;
            atomicAdd(Ab + idx, (float)vre);
            atomicAdd(Ab + idx + 1, (float)vim);
        }
    for (int sv = 0; sv < n_src; ++sv) {
        const size_t qi = (((size_t)b * n_src + sv) * T + orig) * 2;
        const float qre = q[qi], qim = q[qi + 1];
        if (qre == 0.0f && qim == 0.0f) continue;
        float *rb = rhs + ((size_t)b * n_src + sv) * rhs_stride * 2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float sre = (float)sl[i][0], sim = (float)sl[i][1];
            atomicAdd(rb + (size_t)rows[i] * 2, sre * qre - sim * qim);
            atomicAdd(rb + (size_t)rows[i] * 2 + 1, sre * qim + sim * qre);
        }
    }
}

// ---------------------------------------------------------------------------
// Neumann data seen by the field with Robin walls (neumannWithRobin,
// main.swift:1646-1674): qt = q + i k beta (p0 + p1 + p2) / 3 per triangle
// (caller order), i k = (-k_im, k_re); one thread per (triangle, vector),
// vector v of frequency v / vpk.
// ---------------------------------------------------------------------------

__global__ void k_robin_neumann(const int32_t *__restrict__ l2g, int T, int vpk,
                                const float *__restrict__ k_real,
                                const float *__restrict__ k_imag, const float *__restrict__ q,
                                const float *__restrict__ beta, const float *__restrict__ p,
                                int64_t p_stride, float *__restrict__ qt) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int v = blockIdx.y;
    if (t >= T) return;
    const int f = v / vpk;
    const size_t qi = ((size_t)v * T + t) * 2;
    float re = q[qi], im = q[qi + 1];
    const float br = beta[((size_t)f * T + t) * 2], bi = beta[((size_t)f * T + t) * 2 + 1];
    if (br != 0.0f || bi != 0.0f) {
        const float *pv = p + (size_t)v * p_stride * 2;
        float pr = 0.0f, pim = 0.0f;
        for (int l = 0; l < 3; ++l) {
            const int d = l2g[3 * t + l];
            pr += pv[2 * d];
            pim += pv[2 * d + 1];
        }
        pr *= (1.0f / 3.0f);
        pim *= (1.0f / 3.0f);
        const float ikr = k_imag ? -k_imag[f] : 0.0f, iki = k_real[f];
        const float cr = ikr * br - iki * bi, ci = ikr * bi + iki * br;  // i k beta
        re += cr * pr - ci * pim;
        im += cr * pim + ci * pr;
    }
    qt[qi] = re;
    qt[qi + 1] = im;
}

// ---------------------------------------------------------------------------
// Exterior field.  Lanes own observation points; a wave walks a chunk of
// triangles (wave-uniform records).  grid = (obs blocks, triangle chunks, B).
// Chunks of 16 triangles give the few observation waves of a frequency
// enough siblings to hide the records' scalar-load latency (64: 0.63 ms,
// 16: 0.55 ms, 8: 0.56 ms for the 3 x 37 points, 1856 triangles and 40
// frequencies of a quarter-horn sweep).
// ---------------------------------------------------------------------------

constexpr int kFieldChunk = 16;

template <int kMasks>
__global__ __launch_bounds__(256) void k_field(DeviceGeometry g, const float *__restrict__ field_k,
                                               int vpk, const float *__restrict__ p,
                                               int64_t p_stride, const float *__restrict__ q,
                                               int n_obs, const float *__restrict__ obs,
                                               float *__restrict__ out) {
    const int T = g.n_triangles;
    const int o = blockIdx.x * 256 + threadIdx.x;
    // waves wholly past the last observation point have nothing to add (no
    // barrier or LDS below)
    if (blockIdx.x * 256 + (threadIdx.x & ~63) >= n_obs) return;
    const int b = blockIdx.z;  // vector; its wavenumber is field_k[b / vpk]
    const bool ok = o < n_obs;
    const int oo = ok ? o : 0;
    const float ox = obs[oo], oy = obs[n_obs + oo], oz = obs[2 * n_obs + oo];
    const float k = field_k[b / vpk];
    const float *pb = p + (size_t)b * p_stride * 2;
    const float *qb = q + (size_t)b * T * 2;
    float acc_re = 0.f, acc_im = 0.f;
    const int t_begin = blockIdx.y * kFieldChunk;
    const int t_end = min(T, t_begin + kFieldChunk);
    const double ox64 = ox, oy64 = oy, oz64 = oz;
    const double krev64 = (double)k * kInv2PiD;
    for (int t = t_begin; t < t_end; ++t) {
        const float *tf = g.tri_f_aos + (size_t)t * kTriF;
        const int32_t *ti = g.tri_i_aos + (size_t)t * kTriI;
        const float p0r = pb[2 * ti[0]], p0i = pb[2 * ti[0] + 1];
        const float p1r = pb[2 * ti[1]], p1i = pb[2 * ti[1] + 1];
        const float p2r = pb[2 * ti[2]], p2i = pb[2 * ti[2] + 1];
        const float qre = qb[2 * ti[19]], qim = qb[2 * ti[19] + 1];
        const float jac = tf[22];
        // vertex 0 and the two edges in FP64 (the vertices are exact floats)
        double v0[3], e1[3], e2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v0[c] = tf[23 + c];
            e1[c] = (double)tf[26 + c] - v0[c];
            e2[c] = (double)tf[29 + c] - v0[c];
        }
        float tre = 0.f, tim = 0.f;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const float xi = (float)kRule6X[a], eta = (float)kRule6Y[a];
            const float b0 = 1.0f - xi - eta;
            const float pr = b0 * p0r + xi * p1r + eta * p2r;
            const float pi = b0 * p0i + xi * p1i + eta * p2i;
            const float w = (float)kRule6W[a];
            double y[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) y[c] = fma(kRule6X[a], e1[c], fma(kRule6Y[a], e2[c], v0[c]));
#pragma unroll
            for (int m = 0; m < kMasks; ++m) {
                const int mask = g.masks[m];
                const float sx = mask_sign(mask, 1), sy = mask_sign(mask, 2), sz = mask_sign(mask, 4);
                const GD v = kernel_gd_far(sx * y[0] - ox64, sy * y[1] - oy64, sz * y[2] - oz64,
                                           sx * tf[18], sy * tf[19], sz * tf[20], k, krev64);
                // (dG/dn p - G q) w
                const float cre = (v.dre * pr - v.dim * pi) - (v.gre * qre - v.gim * qim);
                const float cim = (v.dre * pi + v.dim * pr) - (v.gre * qim + v.gim * qre);
                tre = fmaf(cre, w, tre);
                tim = fmaf(cim, w, tim);
            }
        }
        acc_re = fmaf(tre, jac, acc_re);
        acc_im = fmaf(tim, jac, acc_im);
    }
    if (ok) {
        atomicAdd(out + ((size_t)b * n_obs + o) * 2, acc_re);
        atomicAdd(out + ((size_t)b * n_obs + o) * 2 + 1, acc_im);
    }
}

// ---------------------------------------------------------------------------
// Tag averages: one block per (tag, frequency), FP64 accumulation.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_tag_average(DeviceGeometry g, const float *__restrict__ p,
                                                     int64_t p_stride, int n_tags,
                                                     const int32_t *__restrict__ tags,
                                                     double *__restrict__ out) {
    const int t_idx = blockIdx.x;
    const int b = blockIdx.y;
    const int tag = tags[t_idx];
    const int T = g.n_triangles;
    const float *pb = p + (size_t)b * p_stride * 2;
    double sr = 0.0, si = 0.0, sa = 0.0;
    for (int t = threadIdx.x; t < T; t += 256) {
        const int32_t *ti = g.tri_i_aos + (size_t)t * kTriI;
        if (ti[18] != tag) continue;
        const double area = (double)g.tri_f_aos[(size_t)t * kTriF + 21];
        if (!(area > 0.0)) continue;
        const double ar = ((double)pb[2 * ti[0]] + (double)pb[2 * ti[1]] + (double)pb[2 * ti[2]]) / 3.0;
        const double ai =
            ((double)pb[2 * ti[0] + 1] + (double)pb[2 * ti[1] + 1] + (double)pb[2 * ti[2] + 1]) / 3.0;
        sr += ar * area;
        si += ai * area;
        sa += area;
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    red[2][threadIdx.x] = sa;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double ta = red[2][0];
        double* o = out + ((size_t)b * n_tags + t_idx) * 2;
        if (ta < 1e-30) {
            o[0] = 0.0;
            o[1] = 0.0;
        } else {
            o[0] = red[0][0] / ta;
            o[1] = red[1][0] / ta;
        }
    }
}

__global__ void k_fill_zero(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_fill_zero_f64(double *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace

void launch_fill_zero(float *ptr, size_t n, hipStream_t stream) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
}

void launch_fill_zero_f64(double *ptr, size_t n, hipStream_t stream) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_zero_f64, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
}

void launch_assemble_regular(const DeviceGeometry &g, int n_batch, const float *k_real,
                             const float *k_imag, const float *q, int n_src, const float *beta,
                             float *A, int64_t lda, int64_t a_stride, float *rhs,
                             int64_t rhs_stride, hipStream_t stream, bool precise) {
    if (g.n_tile_pairs == 0 || n_batch <= 0) return;
    dim3 grid(g.n_tile_pairs, n_batch);
    // two tables of (max tile dofs)^2 complex floats
    const size_t shm = sizeof(float) * 4 * (size_t)g.max_tile_dofs * g.max_tile_dofs;
    auto go = [&](auto complex_k, auto masks) {
        constexpr bool kC = decltype(complex_k)::value;
        constexpr int kM = decltype(masks)::value;
        if (precise)
            hipLaunchKernelGGL((k_assemble_regular<kC, kM, true>), grid, dim3(256), shm, stream, g,
                               k_real, k_imag, q, n_src, beta, A, lda, a_stride, rhs, rhs_stride);
        else
            hipLaunchKernelGGL((k_assemble_regular<kC, kM, false>), grid, dim3(256), shm, stream,
                               g, k_real, k_imag, q, n_src, beta, A, lda, a_stride, rhs,
                               rhs_stride);
    };
    auto by_masks = [&](auto complex_k) {
        switch (g.n_masks) {
        case 1: go(complex_k, std::integral_constant<int, 1>{}); break;
        case 2: go(complex_k, std::integral_constant<int, 2>{}); break;
        default: go(complex_k, std::integral_constant<int, 4>{}); break;
        }
    };
    if (k_imag != nullptr)
        by_masks(std::true_type{});
    else
        by_masks(std::false_type{});
}

void launch_assemble_singular(const DeviceGeometry &g, int n_batch, const float *k_real,
                              const float *k_imag, const float *q, int n_src, const float *beta,
                              float *A, int64_t lda, int64_t a_stride, float *rhs,
                              int64_t rhs_stride, hipStream_t stream) {
    const int waves = g.sg_waves[0] + g.sg_waves[1] + g.sg_waves[2];
    if (waves == 0 || n_batch <= 0) return;
    constexpr int kFB = 2;
    dim3 grid((waves + 3) / 4, (n_batch + kFB - 1) / kFB);
    if (k_imag != nullptr)
        hipLaunchKernelGGL((k_assemble_singular<true, kFB>), grid, dim3(256), 0, stream, g,
                           g.sg_rule, n_batch, k_real, k_imag, q, n_src, beta, A, lda, a_stride, rhs,
                           rhs_stride);
    else
        hipLaunchKernelGGL((k_assemble_singular<false, kFB>), grid, dim3(256), 0, stream, g,
                           g.sg_rule, n_batch, k_real, k_imag, q, n_src, beta, A, lda, a_stride, rhs,
                           rhs_stride);
}

void launch_assemble_multi(const AssemblyJob *jobs_dev, const std::vector<AssemblyJob> &jobs,
                           bool complex_k, hipStream_t stream) {
    if (jobs.empty()) return;
    const int n = (int)jobs.size();
    if (n == 1) {
        // one member: the single-mesh kernels (geometry in the kernel
        // arguments, fewer registers than the member lookup)
        const AssemblyJob &j = jobs[0];
        const float *kim = complex_k ? j.k_imag : nullptr;
        launch_assemble_regular(j.g, j.n_batch, j.k_real, kim, j.q, j.n_src, j.beta, j.A, j.lda,
                                j.a_stride, j.rhs, j.rhs_stride, stream, false);
        launch_assemble_singular(j.g, j.n_batch, j.k_real, kim, j.q, j.n_src, j.beta, j.A, j.lda,
                                 j.a_stride, j.rhs, j.rhs_stride, stream);
        return;
    }
    int reg = 0, sing = 0, dofs = 0;
    for (const AssemblyJob &j : jobs) {
        reg += regular_blocks(j.g, j.n_batch);
        sing += singular_blocks(j.g, j.n_batch);
        dofs = std::max(dofs, j.g.max_tile_dofs);
    }
    const size_t shm = sizeof(float) * 4 * (size_t)dofs * dofs;
    constexpr int kFB = 2;
    auto go = [&](auto complex_k_t) {
        constexpr bool kC = decltype(complex_k_t)::value;
        if (reg > 0) {
            switch (jobs[0].g.n_masks) {
            case 1:
                hipLaunchKernelGGL((k_assemble_regular_multi<kC, 1>), dim3(reg), dim3(256), shm,
                                   stream, jobs_dev, n);
                break;
            case 2:
                hipLaunchKernelGGL((k_assemble_regular_multi<kC, 2>), dim3(reg), dim3(256), shm,
                                   stream, jobs_dev, n);
                break;
            default:
                hipLaunchKernelGGL((k_assemble_regular_multi<kC, 4>), dim3(reg), dim3(256), shm,
                                   stream, jobs_dev, n);
                break;
            }
        }
        if (sing > 0)
            hipLaunchKernelGGL((k_assemble_singular_multi<kC, kFB>), dim3(sing), dim3(256), 0,
                               stream, jobs_dev, n);
    };
    if (complex_k)
        go(std::true_type{});
    else
        go(std::false_type{});
}

void launch_assemble_near(const DeviceGeometry &g, int n_batch, const float *k_real,
                          const float *k_imag, const float *q, int n_src, const float *beta,
                          float *A, int64_t lda, int64_t a_stride, float *rhs,
                          int64_t rhs_stride, hipStream_t stream) {
    if (g.nc_np == 0 || g.nc_rule == nullptr || n_batch <= 0) return;
    dim3 grid((g.nc_np + 255) / 256, n_batch);
    if (k_imag != nullptr)
        hipLaunchKernelGGL((k_assemble_near<true>), grid, dim3(256), 0, stream, g, k_real, k_imag,
                           q, n_src, beta, A, lda, a_stride, rhs, rhs_stride);
    else
        hipLaunchKernelGGL((k_assemble_near<false>), grid, dim3(256), 0, stream, g, k_real, k_imag,
                           q, n_src, beta, A, lda, a_stride, rhs, rhs_stride);
}

void launch_robin_neumann(const DeviceGeometry &g, int n_vec, int vpk, const float *k_real,
                          const float *k_imag, const float *q, const float *beta, const float *p,
                          int64_t p_stride, float *qt, hipStream_t stream) {
    hipLaunchKernelGGL(k_robin_neumann, dim3((g.n_triangles + 255) / 256, n_vec), dim3(256), 0,
                       stream, g.l2g, g.n_triangles, vpk, k_real, k_imag, q, beta, p, p_stride,
                       qt);
}

void launch_field(const DeviceGeometry &g, int n_batch, const float *field_k, int vpk,
                  const float *p, int64_t p_stride, const float *q, int n_obs,
                  const float *obs_3xn, float *out, hipStream_t stream) {
    if (n_obs == 0) return;
    dim3 grid((n_obs + 255) / 256, (g.n_triangles + kFieldChunk - 1) / kFieldChunk, n_batch);
    switch (g.n_masks) {
    case 1:
        hipLaunchKernelGGL((k_field<1>), grid, dim3(256), 0, stream, g, field_k, vpk, p, p_stride, q,
                           n_obs, obs_3xn, out);
        break;
    case 2:
        hipLaunchKernelGGL((k_field<2>), grid, dim3(256), 0, stream, g, field_k, vpk, p, p_stride, q,
                           n_obs, obs_3xn, out);
        break;
    default:
        hipLaunchKernelGGL((k_field<4>), grid, dim3(256), 0, stream, g, field_k, vpk, p, p_stride, q,
                           n_obs, obs_3xn, out);
        break;
    }
}

void launch_tag_average(const DeviceGeometry &g, int n_batch, const float *p, int64_t p_stride,
                        int n_tags, const int32_t *tags_dev, double *out, hipStream_t stream) {
    if (n_tags == 0) return;
    hipLaunchKernelGGL(k_tag_average, dim3(n_tags, n_batch), dim3(256), 0, stream, g, p, p_stride,
                       n_tags, tags_dev, out);
}

}  // namespace hbem

// ---------------------------------------------------------------------------
// Transcendental-pipe probe: the sustained rate of the hardware sine, cosine
// and reciprocal square root (v_sin_f32, v_cos_f32, v_rsq_f32, one third
// each, as in the assembly kernels) over the whole device.  Eight
// independent chains per lane keep the pipes busy; every block runs the same
// count, so ops = blocks * 256 * iters * 8.
// ---------------------------------------------------------------------------

namespace hbem {
namespace {
__global__ __launch_bounds__(256) void k_probe_transcendental(float *out, int iters) {
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = 0.01f * (float)(threadIdx.x + 1) + 0.1f * (float)c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; c += 4) {
            v[c] = __builtin_amdgcn_sinf(v[c]);
            v[c + 1] = __builtin_amdgcn_cosf(v[c + 1]);
            v[c + 2] = __builtin_amdgcn_rsqf(v[c + 2] + 1.0f);
            v[c + 3] = __builtin_amdgcn_sinf(v[c + 3]);
        }
#pragma unroll
        for (int c = 0; c < 8; c += 4) {
            v[c] = __builtin_amdgcn_cosf(v[c]);
            v[c + 1] = __builtin_amdgcn_rsqf(v[c + 1] + 1.0f);
            v[c + 2] = __builtin_amdgcn_sinf(v[c + 2]);
            v[c + 3] = __builtin_amdgcn_cosf(v[c + 3]);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) s += v[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}
}  // namespace

namespace {
__global__ void k_timestamp(long long *slot) { *slot = wall_clock64(); }
}  // namespace

void launch_timestamp(long long *slot, hipStream_t stream) {
    hipLaunchKernelGGL(k_timestamp, dim3(1), dim3(1), 0, stream, slot);
}

void launch_probe_transcendental(float *out, int blocks, int iters, hipStream_t stream) {
    hipLaunchKernelGGL(k_probe_transcendental, dim3(blocks), dim3(256), 0, stream, out, iters);
}
}  // namespace hbem
