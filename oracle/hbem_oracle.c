/*
 * hbem_oracle.c -- CPU restatement of hornlab_metal_bem's per-frequency BEM
 * hot path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of bench.py load this library as the *checker*.  No
 * product code path may call it.
 *
 * Built twice from this one file (see oracle/Makefile):
 *   libhbem_oracle64.so  -- `real` = double (the FP64 oracle)
 *   libhbem_oracle32.so  -- `real` = float  (a port of the reference CPU
 *                           path in its own precision; the CPU baseline)
 *
 * Every routine cites the reference function it restates.  MS = the Swift
 * helper hornlab_metal_bem/metal/native_helper/Sources/HornlabMetalBemNative/
 * main.swift in the reference tree.
 *
 * Geometry arrives exactly as the reference's Metal geometry buffers
 * (hornlab_metal_bem/metal/geometry.py:36-140): float32 SoA vertices
 * [3][nV], int32 plane-major triangles [3][T], int32 P1 local-to-global
 * [T][3], float32 areas [T], float32 plane-major normals [3][T].  The FP64
 * oracle promotes those float32 values, so both precisions see identical
 * inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef HBO_FP32
typedef float real;
#define RSQRT(x) sqrtf(x)
#define RSIN(x) sinf(x)
#define RCOS(x) cosf(x)
#define REXP(x) expf(x)
#else
typedef double real;
#define RSQRT(x) sqrt(x)
#define RSIN(x) sin(x)
#define RCOS(x) cos(x)
#define REXP(x) exp(x)
#endif

#define HBO_EXPORT __attribute__((visibility("default")))

typedef struct {
    real re, im;
} cx;

static inline cx cx_make(real re, real im) {
    cx z;
    z.re = re;
    z.im = im;
    return z;
}
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
static inline cx cx_mul(cx a, cx b) {
    return cx_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cx_scale(cx a, real s) { return cx_make(a.re * s, a.im * s); }

/* Geometry view (layout of geometry.py:36-65). */
typedef struct {
    const real *vertices;     /* [3][n_vertices] */
    const int32_t *triangles; /* [3][n_triangles] */
    const int32_t *l2g;       /* [n_triangles][3] */
    const real *areas;        /* [n_triangles] */
    const real *normals;      /* [3][n_triangles] */
    int32_t n_vertices;
    int32_t n_triangles;
    int32_t n_dofs;
    int32_t symmetry; /* 0 none, 1 yz, 2 xz, 3 yz+xz, 4 xy (MS:703-717) */
} hbo_geom;

static inline int tri_vertex(const hbo_geom *g, int tri, int local) {
    return g->triangles[local * g->n_triangles + tri];
}
static inline int p1_dof(const hbo_geom *g, int tri, int local) {
    return g->l2g[tri * 3 + local];
}
static inline real vx(const hbo_geom *g, int v, int c) {
    return (real)g->vertices[c * g->n_vertices + v];
}
static inline real tri_normal(const hbo_geom *g, int tri, int c) {
    return (real)g->normals[c * g->n_triangles + tri];
}

/* 2^planes row weight: MS:719-734 (symmetryRowWeight). */
static real row_weight(const hbo_geom *g) {
    real w = 1;
    if (g->symmetry & 1) w *= 2;
    if (g->symmetry & 2) w *= 2;
    if (g->symmetry & 4) w *= 2;
    return w;
}

/* Image masks of a symmetry code: MS:1570-1584 (mirrorMasks). */
static int image_masks(int symmetry, int *masks) {
    switch (symmetry) {
    case 1: masks[0] = 1; return 1;
    case 2: masks[0] = 2; return 1;
    case 4: masks[0] = 4; return 1;
    case 3: masks[0] = 1; masks[1] = 2; masks[2] = 3; return 3;
    default: return 0;
    }
}

/* ---------------------------------------------------------------------- */
/* Quadrature rules                                                        */
/* ---------------------------------------------------------------------- */

/* Degree-4 six-point triangle rule with weights summing to 1/2:
 * MS:1472-1489 (triangleRule6).  Point/weight values are the published
 * Strang-Fix/Dunavant degree-4 rule; they are written as in the reference
 * so that both precisions round the same decimal strings. */
static const double RULE6_X[6] = {0.4459484909159651, 0.0915762135097710,
                                  0.1081030181680700, 0.4459484909159651,
                                  0.8168475729804590, 0.0915762135097710};
static const double RULE6_Y[6] = {0.4459484909159651, 0.0915762135097700,
                                  0.4459484909159651, 0.1081030181680700,
                                  0.0915762135097700, 0.8168475729804580};
static const double RULE6_W[6] = {0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220,
                                  0.5 * 0.2233815896780110, 0.5 * 0.2233815896780110,
                                  0.5 * 0.1099517436553220, 0.5 * 0.1099517436553220};

HBO_EXPORT void hbo_rule6(real *qx, real *qy, real *qw) {
    for (int i = 0; i < 6; ++i) {
        qx[i] = (real)RULE6_X[i];
        qy[i] = (real)RULE6_Y[i];
        qw[i] = (real)RULE6_W[i];
    }
}

/* Gauss-Legendre on [0,1].  Order 4 uses the closed form of MS:2626-2635
 * (gaussRule1D4); other orders (used only to study convergence) use Newton
 * iteration on the Legendre recurrence. */
static void gauss01(int order, real *x, real *w) {
    if (order == 4) {
        real s65 = RSQRT((real)6.0 / (real)5.0);
        real x1 = RSQRT((real)3.0 / (real)7.0 - (real)2.0 / (real)7.0 * s65) / 2;
        real x2 = RSQRT((real)3.0 / (real)7.0 + (real)2.0 / (real)7.0 * s65) / 2;
        real s30 = RSQRT((real)30.0);
        real w1 = ((real)18.0 + s30) / 72;
        real w2 = ((real)18.0 - s30) / 72;
        x[0] = (real)0.5 - x2; x[1] = (real)0.5 - x1;
        x[2] = (real)0.5 + x1; x[3] = (real)0.5 + x2;
        w[0] = w2; w[1] = w1; w[2] = w1; w[3] = w2;
        return;
    }
    double xd[64], wd[64];
    for (int i = 0; i < order; ++i) {
        double z = cos(M_PI * (i + 0.75) / (order + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= order; ++j) {
                double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            double pn = (order == 1) ? z : p1;
            double pm = (order == 1) ? 1.0 : p0;
            dp = order * (z * pn - pm) / (z * z - 1.0);
            double dz = pn / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        xd[i] = 0.5 * (1.0 + z);
        wd[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
    /* the Newton starts run from +1 downwards: store ascending */
    for (int i = 0; i < order; ++i) {
        x[i] = (real)xd[order - 1 - i];
        w[i] = (real)wd[order - 1 - i];
    }
}

/* One Sauter-Schwab region point, stored in the reference triangle's
 * (x - y, y) form: MS:2637-2650 (appendDuffy). */
static int put_duffy(real *tx, real *ty, real *sx, real *sy, real *w, int n,
                     real a1, real a2, real b1, real b2, real weight) {
    tx[n] = a1 - a2;
    ty[n] = a2;
    sx[n] = b1 - b2;
    sy[n] = b2;
    w[n] = weight;
    return n + 1;
}

/* Number of points of the singular rule of `kind` for a 1D order. */
HBO_EXPORT int hbo_duffy_rule_size(int kind, int order) {
    int o4 = order * order * order * order;
    if (kind == 1) return 6 * o4;
    if (kind == 2) return 5 * o4;
    if (kind == 3) return 2 * o4;
    return 0;
}

/* Sauter-Schwab singular rules for coincident (1), edge (2) and vertex (3)
 * pairs: MS:2652-2699 (duffyRule).  The region maps and Jacobian factors
 * are the published Sauter-Schwab construction; the loop order (a, b, c, d)
 * with xi = x[b], eta1 = x[a], eta2 = x[c], eta3 = x[d] follows the
 * reference so the point sequence is the same. */
HBO_EXPORT int hbo_duffy_rule(int kind, int order, real *tx, real *ty, real *sx, real *sy,
                              real *w) {
    real xs[64], ws[64];
    if (order < 1 || order > 64) return -1;
    gauss01(order, xs, ws);
    int n = 0;
    for (int a = 0; a < order; ++a)
        for (int b = 0; b < order; ++b)
            for (int c = 0; c < order; ++c)
                for (int d = 0; d < order; ++d) {
                    real xi = xs[b], e1 = xs[a], e2 = xs[c], e3 = xs[d];
                    real e12 = e1 * e2, e123 = e12 * e3;
                    real base = ws[a] * ws[b] * ws[c] * ws[d];
                    if (kind == 1) {
                        real wt = base * xi * xi * xi * e1 * e1 * e2;
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * (1 - e1 + e12),
                                      xi * (1 - e123), xi * (1 - e1), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e123), xi * (1 - e1),
                                      xi, xi * (1 - e1 + e12), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * (e1 - e12 + e123),
                                      xi * (1 - e12), xi * (e1 - e12), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e12), xi * (e1 - e12),
                                      xi, xi * (e1 - e12 + e123), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e123), xi * (e1 - e123),
                                      xi, xi * (e1 - e12), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * (e1 - e12),
                                      xi * (1 - e123), xi * (e1 - e123), wt);
                    } else if (kind == 2) {
                        real wt = base * xi * xi * xi * e1 * e1;
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * e1 * e3, xi * (1 - e12),
                                      xi * e1 * (1 - e2), wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * e1, xi * (1 - e123),
                                      xi * e1 * e2 * (1 - e3), wt * e2);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e12), xi * e1 * (1 - e2),
                                      xi, xi * e123, wt * e2);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e123), xi * e12 * (1 - e3),
                                      xi, xi * e1, wt * e2);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * (1 - e123),
                                      xi * e1 * (1 - e2 * e3), xi, xi * e12, wt * e2);
                    } else if (kind == 3) {
                        real wt = base * xi * xi * xi * e2;
                        n = put_duffy(tx, ty, sx, sy, w, n, xi, xi * e1, xi * e2, xi * e2 * e3, wt);
                        n = put_duffy(tx, ty, sx, sy, w, n, xi * e2, xi * e2 * e3, xi, xi * e1, wt);
                    } else {
                        return -1;
                    }
                }
    return n;
}

/* Map a rule point onto the pair's shared-vertex configuration:
 * MS:2701-2735 (refX/refY/remapSingular). */
static void remap_singular(real x, real y, int kind, int l1, int l2, real *ox, real *oy) {
    if (kind == 1) {
        *ox = x;
        *oy = y;
        return;
    }
    if (kind == 2) {
        int vc = 3 - l1 - l2;
        real ax = (l1 == 1), ay = (l1 == 2);
        real bx = (l2 == 1), by = (l2 == 2);
        real cxr = (vc == 1), cyr = (vc == 2);
        *ox = ax + x * (bx - ax) + y * (cxr - ax);
        *oy = ay + x * (by - ay) + y * (cyr - ay);
        return;
    }
    if (l1 == 0) {
        *ox = x;
        *oy = y;
    } else if (l1 == 1) {
        *ox = 1 - x - y;
        *oy = y;
    } else {
        *ox = x;
        *oy = 1 - x - y;
    }
}

/* ---------------------------------------------------------------------- */
/* Kernels and geometry helpers                                            */
/* ---------------------------------------------------------------------- */

#define INV4PI ((real)0.07957747154594767)

/* G and dG/dn_y for r = y - x: MS:1513-1568 (helmholtzGComplex,
 * helmholtzDlpComplex): G = e^{i k r} e^{-k_im r} / (4 pi r);
 * dG/dn_y = G (i k - k_im - 1/r) (r . n_y)/r.  Zero for r^2 <= 0. */
static inline void kern_g_dlp(real dx, real dy, real dz, real nx, real ny, real nz, real k,
                              real kim, cx *g, cx *d) {
    real r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= 0) {
        *g = cx_make(0, 0);
        *d = cx_make(0, 0);
        return;
    }
    real r = RSQRT(r2);
    real phase = k * r;
    real scale = REXP(-kim * r) * INV4PI / r;
    real gre = RCOS(phase) * scale;
    real gim = RSIN(phase) * scale;
    real proj = (dx * nx + dy * ny + dz * nz) / r;
    real fre = -1 / r - kim;
    real fim = k;
    *g = cx_make(gre, gim);
    *d = cx_make((gre * fre - gim * fim) * proj, (gre * fim + gim * fre) * proj);
}

HBO_EXPORT void hbo_kernel(real dx, real dy, real dz, real nx, real ny, real nz, real k, real kim,
                           real *out4) {
    cx g, d;
    kern_g_dlp(dx, dy, dz, nx, ny, nz, k, kim, &g, &d);
    out4[0] = g.re;
    out4[1] = g.im;
    out4[2] = d.re;
    out4[3] = d.im;
}

/* x(xi, eta) = b1 v1 + b2 v2 + b3 v3: MS:1495-1507 (pointOnTriangle). */
static inline void point_on_tri(const hbo_geom *g, int tri, real xi, real eta, real *p) {
    real b1 = 1 - xi - eta, b2 = xi, b3 = eta;
    int i1 = tri_vertex(g, tri, 0), i2 = tri_vertex(g, tri, 1), i3 = tri_vertex(g, tri, 2);
    for (int c = 0; c < 3; ++c) p[c] = b1 * vx(g, i1, c) + b2 * vx(g, i2, c) + b3 * vx(g, i3, c);
}

static inline void mirror3(real *p, int mask) {
    if (mask & 1) p[0] = -p[0];
    if (mask & 2) p[1] = -p[1];
    if (mask & 4) p[2] = -p[2];
}

/* Odd masks reverse the image triangle's local order (MS:2118-2150). */
static inline int mask_is_odd(int mask) {
    int b = 0;
    while (mask) {
        b += mask & 1;
        mask >>= 1;
    }
    return b & 1;
}
static inline int image_local(int local, int mask) {
    if (!mask_is_odd(mask)) return local;
    return local == 1 ? 2 : (local == 2 ? 1 : local);
}
static inline void image_ref_to_original(real x, real y, int mask, real *ox, real *oy) {
    if (mask_is_odd(mask)) {
        *ox = y;
        *oy = x;
    } else {
        *ox = x;
        *oy = y;
    }
}

/* ---------------------------------------------------------------------- */
/* Regular assembly: MS:1726-1841 (assembleRegularReference)               */
/* ---------------------------------------------------------------------- */

/* Greedy vertex-disjoint colouring of the triangles, so that one colour's
 * test triangles write disjoint matrix rows and can run in parallel. */
static int color_triangles(const hbo_geom *g, int *color) {
    int T = g->n_triangles, nv = g->n_vertices;
    int32_t *vstamp = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv * 64);
    /* vstamp[v*64 + c] = 1 if colour c already touches vertex v */
    memset(vstamp, 0, sizeof(int32_t) * (size_t)nv * 64);
    int ncolors = 0;
    for (int t = 0; t < T; ++t) {
        int c = 0;
        for (; c < 64; ++c) {
            int ok = 1;
            for (int l = 0; l < 3; ++l)
                if (vstamp[(size_t)tri_vertex(g, t, l) * 64 + c]) ok = 0;
            if (ok) break;
        }
        if (c == 64) {
            free(vstamp);
            return -1;
        }
        color[t] = c;
        for (int l = 0; l < 3; ++l) vstamp[(size_t)tri_vertex(g, t, l) * 64 + c] = 1;
        if (c + 1 > ncolors) ncolors = c + 1;
    }
    free(vstamp);
    return ncolors;
}

/* A (row-major n x n) and rhs are ADDED to; beta (optional, per triangle)
 * is the Robin admittance folded as in MS:1745-1820. */
static void assemble_regular(const hbo_geom *g, const unsigned char *sel, real k, real kim,
                             const real *q_re, const real *q_im, const real *beta_re,
                             const real *beta_im, real *A_re, real *A_im, real *rhs_re,
                             real *rhs_im) {
    int T = g->n_triangles, n = g->n_dofs;
    int masks[3];
    int nm = image_masks(g->symmetry, masks);
    real rw = row_weight(g);
    real qx[6], qy[6], qw[6];
    hbo_rule6(qx, qy, qw);
    cx iK = cx_make(-kim, k);
    int *color = (int *)malloc(sizeof(int) * (size_t)T);
    int ncolors = color_triangles(g, color);
    int *order = (int *)malloc(sizeof(int) * (size_t)T);
    int *start = (int *)calloc((size_t)ncolors + 1, sizeof(int));
    for (int t = 0; t < T; ++t) start[color[t] + 1]++;
    for (int c = 0; c < ncolors; ++c) start[c + 1] += start[c];
    {
        int *cur = (int *)malloc(sizeof(int) * (size_t)ncolors);
        for (int c = 0; c < ncolors; ++c) cur[c] = start[c];
        for (int t = 0; t < T; ++t) order[cur[color[t]]++] = t;
        free(cur);
    }
    for (int c = 0; c < ncolors; ++c) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int oi = start[c]; oi < start[c + 1]; ++oi) {
            int test = order[oi];
            real tp[6][3];
            for (int a = 0; a < 6; ++a) point_on_tri(g, test, qx[a], qy[a], tp[a]);
            for (int trial = 0; trial < T; ++trial) {
                if (sel && !sel[test] && !sel[trial]) continue;
                real nnx = tri_normal(g, trial, 0), nny = tri_normal(g, trial, 1),
                     nnz = tri_normal(g, trial, 2);
                real jac = (2 * (real)g->areas[test]) * (2 * (real)g->areas[trial]);
                cx block[9], slp[3];
                for (int i = 0; i < 9; ++i) block[i] = cx_make(0, 0);
                for (int i = 0; i < 3; ++i) slp[i] = cx_make(0, 0);
                for (int a = 0; a < 6; ++a) {
                    real tb[3] = {1 - qx[a] - qy[a], qx[a], qy[a]};
                    for (int b = 0; b < 6; ++b) {
                        real sp[3];
                        point_on_tri(g, trial, qx[b], qy[b], sp);
                        real sb[3] = {1 - qx[b] - qy[b], qx[b], qy[b]};
                        real w = qw[a] * qw[b] * jac;
                        cx gg, dd;
                        kern_g_dlp(sp[0] - tp[a][0], sp[1] - tp[a][1], sp[2] - tp[a][2], nnx, nny,
                                   nnz, k, kim, &gg, &dd);
                        for (int i = 0; i < 3; ++i) {
                            slp[i] = cx_add(slp[i], cx_scale(gg, tb[i] * w));
                            for (int j = 0; j < 3; ++j)
                                block[i * 3 + j] =
                                    cx_add(block[i * 3 + j], cx_scale(dd, tb[i] * sb[j] * w));
                        }
                        for (int m = 0; m < nm; ++m) {
                            real ip[3] = {sp[0], sp[1], sp[2]};
                            real in[3] = {nnx, nny, nnz};
                            mirror3(ip, masks[m]);
                            mirror3(in, masks[m]);
                            kern_g_dlp(ip[0] - tp[a][0], ip[1] - tp[a][1], ip[2] - tp[a][2], in[0],
                                       in[1], in[2], k, kim, &gg, &dd);
                            for (int i = 0; i < 3; ++i) {
                                slp[i] = cx_add(slp[i], cx_scale(gg, tb[i] * w));
                                for (int j = 0; j < 3; ++j)
                                    block[i * 3 + j] = cx_add(block[i * 3 + j],
                                                              cx_scale(dd, tb[i] * sb[j] * w));
                            }
                        }
                    }
                }
                cx qt = cx_make(q_re ? q_re[trial] : 0, q_im ? q_im[trial] : 0);
                cx beta = cx_make(beta_re ? beta_re[trial] : 0, beta_im ? beta_im[trial] : 0);
                int has_robin = (beta.re != 0 || beta.im != 0);
                cx coupling = cx_mul(iK, beta);
                for (int i = 0; i < 3; ++i) {
                    int row = p1_dof(g, test, i);
                    cx r = cx_scale(cx_mul(slp[i], qt), rw);
                    rhs_re[row] += r.re;
                    rhs_im[row] += r.im;
                    for (int j = 0; j < 3; ++j) {
                        int col = p1_dof(g, trial, j);
                        cx term = block[i * 3 + j];
                        if (has_robin)
                            term = cx_sub(term, cx_scale(cx_mul(slp[i], coupling), (real)(1.0 / 3.0)));
                        A_re[(size_t)row * n + col] += term.re * rw;
                        A_im[(size_t)row * n + col] += term.im * rw;
                    }
                }
                if (test == trial) {
                    for (int i = 0; i < 3; ++i) {
                        int row = p1_dof(g, test, i);
                        for (int j = 0; j < 3; ++j) {
                            int col = p1_dof(g, trial, j);
                            real mass = (real)g->areas[test] *
                                        (i == j ? (real)(1.0 / 6.0) : (real)(1.0 / 12.0));
                            A_re[(size_t)row * n + col] -= (real)0.5 * mass * rw;
                        }
                    }
                }
            }
        }
    }
    free(order);
    free(start);
    free(color);
}

HBO_EXPORT void hbo_assemble_regular(const hbo_geom *g, real k, real kim, const real *q_re,
                                     const real *q_im, const real *beta_re, const real *beta_im,
                                     real *A_re, real *A_im, real *rhs_re, real *rhs_im) {
    assemble_regular(g, NULL, k, kim, q_re, q_im, beta_re, beta_im, A_re, A_im, rhs_re, rhs_im);
}

/* ---------------------------------------------------------------------- */
/* Singular pair list: MS:2107-2387                                        */
/* ---------------------------------------------------------------------- */

/* Coordinate key with 1e-6 quantisation: MS:2107-2112 (coordinateKey);
 * Swift's rounded() is round-half-away-from-zero, as C's llround. */
typedef struct {
    int64_t x, y, z;
} vkey;

static vkey key_of(real x, real y, real z) {
    vkey k;
    k.x = llround((double)x / 1.0e-6);
    k.y = llround((double)y / 1.0e-6);
    k.z = llround((double)z / 1.0e-6);
    return k;
}
static int key_eq(vkey a, vkey b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

static vkey vertex_key_masked(const hbo_geom *g, int v, int mask) {
    real p[3] = {vx(g, v, 0), vx(g, v, 1), vx(g, v, 2)};
    mirror3(p, mask);
    return key_of(p[0], p[1], p[2]);
}

typedef struct {
    int count, tl1, tl2, sl1, sl2;
} shared_ids;

/* MS:2083-2105 (sharedLocalIds): shared vertices by index. */
static shared_ids shared_by_index(const hbo_geom *g, int test, int trial) {
    shared_ids s = {0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        int tv = tri_vertex(g, test, i);
        for (int j = 0; j < 3; ++j)
            if (tv == tri_vertex(g, trial, j)) {
                s.count++;
                if (s.count == 1) {
                    s.tl1 = i;
                    s.sl1 = j;
                } else {
                    s.tl2 = i;
                    s.sl2 = j;
                }
            }
    }
    return s;
}

/* MS:2152-2193 (imageSharedLocalIds): shared vertices of the mirrored
 * triangles by coordinate key, reported in image-local order. */
static shared_ids shared_by_key(const hbo_geom *g, int test, int trial, int tm, int sm) {
    if (tm == 0 && sm == 0) return shared_by_index(g, test, trial);
    shared_ids s = {0, 0, 0, 0, 0};
    for (int tfl = 0; tfl < 3; ++tfl) {
        int i = image_local(tfl, tm);
        vkey tk = vertex_key_masked(g, tri_vertex(g, test, i), tm);
        for (int sfl = 0; sfl < 3; ++sfl) {
            int j = image_local(sfl, sm);
            vkey sk = vertex_key_masked(g, tri_vertex(g, trial, j), sm);
            if (key_eq(tk, sk)) {
                s.count++;
                if (s.count == 1) {
                    s.tl1 = tfl;
                    s.sl1 = sfl;
                } else {
                    s.tl2 = tfl;
                    s.sl2 = sfl;
                }
            }
        }
    }
    return s;
}

/* Pair record: test, trial, test mask, trial mask, kind, tl1, tl2, sl1, sl2. */
#define PAIR_W 9

static int append_pair(int32_t *out, int cap, int n, int test, int trial, int tm, int sm,
                       shared_ids s, int64_t *counts) {
    int kind;
    if (s.count == 3) {
        kind = 1;
        s.tl1 = 0;
        s.tl2 = 1;
        s.sl1 = 0;
        s.sl2 = 1;
    } else if (s.count == 2) {
        kind = 2;
    } else if (s.count == 1) {
        kind = 3;
    } else {
        return n;
    }
    counts[kind - 1]++;
    if (out && n < cap) {
        int32_t *p = out + (size_t)n * PAIR_W;
        p[0] = test;
        p[1] = trial;
        p[2] = tm;
        p[3] = sm;
        p[4] = kind;
        p[5] = s.tl1;
        p[6] = s.tl2;
        p[7] = s.sl1;
        p[8] = s.sl2;
    }
    return n + 1;
}

/* Triangles incident to each vertex id (CSR). */
typedef struct {
    int *start, *tris;
} csr_t;

static csr_t build_csr(int nkeys, int T, const int *ids /* [T*3] */) {
    csr_t c;
    c.start = (int *)calloc((size_t)nkeys + 1, sizeof(int));
    c.tris = (int *)malloc(sizeof(int) * (size_t)T * 3);
    for (int t = 0; t < T; ++t)
        for (int l = 0; l < 3; ++l) c.start[ids[t * 3 + l] + 1]++;
    for (int k = 0; k < nkeys; ++k) c.start[k + 1] += c.start[k];
    int *cur = (int *)malloc(sizeof(int) * (size_t)nkeys);
    for (int k = 0; k < nkeys; ++k) cur[k] = c.start[k];
    /* ascending triangle order within each vertex, as the reference appends */
    for (int t = 0; t < T; ++t)
        for (int l = 0; l < 3; ++l) c.tris[cur[ids[t * 3 + l]]++] = t;
    free(cur);
    return c;
}

static int cmp_key(const void *a, const void *b) {
    const vkey *x = (const vkey *)a, *y = (const vkey *)b;
    if (x->x != y->x) return x->x < y->x ? -1 : 1;
    if (x->y != y->y) return x->y < y->y ? -1 : 1;
    if (x->z != y->z) return x->z < y->z ? -1 : 1;
    return 0;
}

static int key_index(const vkey *keys, int n, vkey k) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        int c = cmp_key(&keys[mid], &k);
        if (c == 0) return mid;
        if (c < 0)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

/* Builds the singular pair list.  out may be NULL to count only.  Returns the
 * pair count; counts[0..2] = coincident, edge, vertex.  Non-symmetric meshes
 * follow MS:2256-2299 (buildRealDuffyPairList, vertex-index sharing);
 * symmetric ones MS:2301-2380 (buildSymmetryDuffyPairList, coordinate keys
 * over every (trial mask, test mask) combination). */
HBO_EXPORT int hbo_duffy_pairs(const hbo_geom *g, int32_t *out, int cap, int64_t *counts) {
    int T = g->n_triangles;
    counts[0] = counts[1] = counts[2] = 0;
    int n = 0;
    int *stamp = (int *)calloc((size_t)T, sizeof(int));
    int *cand = (int *)malloc(sizeof(int) * (size_t)T);
    int st = 0;
    if (g->symmetry == 0) {
        int *ids = (int *)malloc(sizeof(int) * (size_t)T * 3);
        for (int t = 0; t < T; ++t)
            for (int l = 0; l < 3; ++l) ids[t * 3 + l] = tri_vertex(g, t, l);
        csr_t c = build_csr(g->n_vertices, T, ids);
        for (int trial = 0; trial < T; ++trial) {
            st++;
            int nc = 0;
            for (int l = 0; l < 3; ++l) {
                int v = tri_vertex(g, trial, l);
                for (int p = c.start[v]; p < c.start[v + 1]; ++p) {
                    int test = c.tris[p];
                    if (stamp[test] != st) {
                        stamp[test] = st;
                        cand[nc++] = test;
                    }
                }
            }
            for (int ci = 0; ci < nc; ++ci)
                n = append_pair(out, cap, n, cand[ci], trial, 0, 0,
                                shared_by_index(g, cand[ci], trial), counts);
        }
        free(ids);
        free(c.start);
        free(c.tris);
    } else {
        int masks[4] = {0, 0, 0, 0};
        int nm = 1 + image_masks(g->symmetry, masks + 1);
        int nv = g->n_vertices;
        vkey *all = (vkey *)malloc(sizeof(vkey) * (size_t)nv * nm);
        for (int m = 0; m < nm; ++m)
            for (int v = 0; v < nv; ++v) all[(size_t)m * nv + v] = vertex_key_masked(g, v, masks[m]);
        vkey *uniq = (vkey *)malloc(sizeof(vkey) * (size_t)nv * nm);
        memcpy(uniq, all, sizeof(vkey) * (size_t)nv * nm);
        qsort(uniq, (size_t)nv * nm, sizeof(vkey), cmp_key);
        int nu = 0;
        for (int i = 0; i < nv * nm; ++i)
            if (nu == 0 || !key_eq(uniq[nu - 1], uniq[i])) uniq[nu++] = uniq[i];
        int *idm = (int *)malloc(sizeof(int) * (size_t)nv * nm); /* id of vertex v under mask m */
        for (int i = 0; i < nv * nm; ++i) idm[i] = key_index(uniq, nu, all[i]);
        csr_t *by_mask = (csr_t *)malloc(sizeof(csr_t) * (size_t)nm);
        int *ids = (int *)malloc(sizeof(int) * (size_t)T * 3);
        for (int m = 0; m < nm; ++m) {
            for (int t = 0; t < T; ++t)
                for (int l = 0; l < 3; ++l) ids[t * 3 + l] = idm[(size_t)m * nv + tri_vertex(g, t, l)];
            by_mask[m] = build_csr(nu, T, ids);
        }
        for (int trial = 0; trial < T; ++trial)
            for (int smi = 0; smi < nm; ++smi)
                for (int tmi = 0; tmi < nm; ++tmi) {
                    st++;
                    int nc = 0;
                    for (int l = 0; l < 3; ++l) {
                        int key = idm[(size_t)smi * nv + tri_vertex(g, trial, l)];
                        csr_t *c = &by_mask[tmi];
                        for (int p = c->start[key]; p < c->start[key + 1]; ++p) {
                            int test = c->tris[p];
                            if (stamp[test] != st) {
                                stamp[test] = st;
                                cand[nc++] = test;
                            }
                        }
                    }
                    for (int ci = 0; ci < nc; ++ci)
                        n = append_pair(out, cap, n, cand[ci], trial, masks[tmi], masks[smi],
                                        shared_by_key(g, cand[ci], trial, masks[tmi], masks[smi]),
                                        counts);
                }
        for (int m = 0; m < nm; ++m) {
            free(by_mask[m].start);
            free(by_mask[m].tris);
        }
        free(by_mask);
        free(ids);
        free(idm);
        free(uniq);
        free(all);
    }
    free(stamp);
    free(cand);
    return n;
}

/* ---------------------------------------------------------------------- */
/* Pair blocks: MS:2788-2972                                               */
/* ---------------------------------------------------------------------- */

/* Regular 6x6 block of one (test, trial, masks) pair with CPU semantics
 * (self point dropped by the r^2 <= 0 guard): MS:2788-2838. */
static void regular_pair_block(const hbo_geom *g, int test, int trial, int tm, int sm, real k,
                               real kim, cx *slp, cx *dlp) {
    real qx[6], qy[6], qw[6];
    hbo_rule6(qx, qy, qw);
    real nrm[3] = {tri_normal(g, trial, 0), tri_normal(g, trial, 1), tri_normal(g, trial, 2)};
    mirror3(nrm, sm);
    real jac = (2 * (real)g->areas[test]) * (2 * (real)g->areas[trial]);
    for (int i = 0; i < 3; ++i) slp[i] = cx_make(0, 0);
    for (int i = 0; i < 9; ++i) dlp[i] = cx_make(0, 0);
    for (int a = 0; a < 6; ++a) {
        real txi, teta, tp[3];
        image_ref_to_original(qx[a], qy[a], tm, &txi, &teta);
        point_on_tri(g, test, txi, teta, tp);
        mirror3(tp, tm);
        real tb[3] = {1 - txi - teta, txi, teta};
        for (int b = 0; b < 6; ++b) {
            real sxi, seta, sp[3];
            image_ref_to_original(qx[b], qy[b], sm, &sxi, &seta);
            point_on_tri(g, trial, sxi, seta, sp);
            mirror3(sp, sm);
            real sb[3] = {1 - sxi - seta, sxi, seta};
            cx gg, dd;
            kern_g_dlp(sp[0] - tp[0], sp[1] - tp[1], sp[2] - tp[2], nrm[0], nrm[1], nrm[2], k, kim,
                       &gg, &dd);
            real w = qw[a] * qw[b] * jac;
            for (int i = 0; i < 3; ++i) {
                slp[i] = cx_add(slp[i], cx_scale(gg, tb[i] * w));
                for (int j = 0; j < 3; ++j)
                    dlp[i * 3 + j] = cx_add(dlp[i * 3 + j], cx_scale(dd, tb[i] * sb[j] * w));
            }
        }
    }
}

/* Subdivided block (4^level sub-triangles per side, 6-point rule on each):
 * MS:2758-2786 (referenceSubtriangles) and MS:2840-2910. */
static void subdivided_pair_block(const hbo_geom *g, int test, int trial, int tm, int sm, real k,
                                  real kim, int level, cx *slp, cx *dlp) {
    real qx[6], qy[6], qw[6];
    hbo_rule6(qx, qy, qw);
    int ns = 1;
    for (int l = 0; l < level; ++l) ns *= 4;
    real *sub = (real *)malloc(sizeof(real) * 6 * (size_t)ns);
    sub[0] = 0; sub[1] = 0; sub[2] = 1; sub[3] = 0; sub[4] = 0; sub[5] = 1;
    int cur = 1;
    for (int l = 0; l < level; ++l) {
        real *next = (real *)malloc(sizeof(real) * 6 * (size_t)cur * 4);
        for (int s = 0; s < cur; ++s) {
            real *t = sub + 6 * s;
            real ax = t[0], ay = t[1], bx = t[2], by = t[3], cxv = t[4], cyv = t[5];
            real abx = (ax + bx) * (real)0.5, aby = (ay + by) * (real)0.5;
            real bcx = (bx + cxv) * (real)0.5, bcy = (by + cyv) * (real)0.5;
            real cax = (cxv + ax) * (real)0.5, cay = (cyv + ay) * (real)0.5;
            real v[4][6] = {{ax, ay, abx, aby, cax, cay},
                            {abx, aby, bx, by, bcx, bcy},
                            {cax, cay, bcx, bcy, cxv, cyv},
                            {abx, aby, bcx, bcy, cax, cay}};
            for (int q = 0; q < 4; ++q) memcpy(next + 6 * (4 * s + q), v[q], sizeof(real) * 6);
        }
        free(sub);
        sub = next;
        cur *= 4;
    }
    real nrm[3] = {tri_normal(g, trial, 0), tri_normal(g, trial, 1), tri_normal(g, trial, 2)};
    mirror3(nrm, sm);
    real tjac = 2 * (real)g->areas[test], sjac = 2 * (real)g->areas[trial];
    for (int i = 0; i < 3; ++i) slp[i] = cx_make(0, 0);
    for (int i = 0; i < 9; ++i) dlp[i] = cx_make(0, 0);
    for (int ts = 0; ts < ns; ++ts) {
        real *T1 = sub + 6 * ts;
        real tdet = (real)fabs((double)((T1[2] - T1[0]) * (T1[5] - T1[1]) - (T1[3] - T1[1]) * (T1[4] - T1[0])));
        real tsubjac = tjac * tdet;
        for (int ss = 0; ss < ns; ++ss) {
            real *S1 = sub + 6 * ss;
            real sdet = (real)fabs((double)((S1[2] - S1[0]) * (S1[5] - S1[1]) - (S1[3] - S1[1]) * (S1[4] - S1[0])));
            real jac = tsubjac * sjac * sdet;
            for (int a = 0; a < 6; ++a) {
                real irx = T1[0] + qx[a] * (T1[2] - T1[0]) + qy[a] * (T1[4] - T1[0]);
                real iry = T1[1] + qx[a] * (T1[3] - T1[1]) + qy[a] * (T1[5] - T1[1]);
                real txi, teta, tp[3];
                image_ref_to_original(irx, iry, tm, &txi, &teta);
                point_on_tri(g, test, txi, teta, tp);
                mirror3(tp, tm);
                real tb[3] = {1 - txi - teta, txi, teta};
                for (int b = 0; b < 6; ++b) {
                    real jrx = S1[0] + qx[b] * (S1[2] - S1[0]) + qy[b] * (S1[4] - S1[0]);
                    real jry = S1[1] + qx[b] * (S1[3] - S1[1]) + qy[b] * (S1[5] - S1[1]);
                    real sxi, seta, sp[3];
                    image_ref_to_original(jrx, jry, sm, &sxi, &seta);
                    point_on_tri(g, trial, sxi, seta, sp);
                    mirror3(sp, sm);
                    real sb[3] = {1 - sxi - seta, sxi, seta};
                    cx gg, dd;
                    kern_g_dlp(sp[0] - tp[0], sp[1] - tp[1], sp[2] - tp[2], nrm[0], nrm[1], nrm[2],
                               k, kim, &gg, &dd);
                    real w = qw[a] * qw[b] * jac;
                    for (int i = 0; i < 3; ++i) {
                        slp[i] = cx_add(slp[i], cx_scale(gg, tb[i] * w));
                        for (int j = 0; j < 3; ++j)
                            dlp[i * 3 + j] = cx_add(dlp[i * 3 + j], cx_scale(dd, tb[i] * sb[j] * w));
                    }
                }
            }
        }
    }
    free(sub);
}

typedef struct {
    int order;
    int n[3];
    real *tx[3], *ty[3], *sx[3], *sy[3], *w[3];
} duffy_rules;

static void rules_make(duffy_rules *r, int order) {
    r->order = order;
    for (int k = 0; k < 3; ++k) {
        int m = hbo_duffy_rule_size(k + 1, order);
        r->tx[k] = (real *)malloc(sizeof(real) * (size_t)m);
        r->ty[k] = (real *)malloc(sizeof(real) * (size_t)m);
        r->sx[k] = (real *)malloc(sizeof(real) * (size_t)m);
        r->sy[k] = (real *)malloc(sizeof(real) * (size_t)m);
        r->w[k] = (real *)malloc(sizeof(real) * (size_t)m);
        r->n[k] = hbo_duffy_rule(k + 1, order, r->tx[k], r->ty[k], r->sx[k], r->sy[k], r->w[k]);
    }
}
static void rules_free(duffy_rules *r) {
    for (int k = 0; k < 3; ++k) {
        free(r->tx[k]);
        free(r->ty[k]);
        free(r->sx[k]);
        free(r->sy[k]);
        free(r->w[k]);
    }
}

/* Singular (Sauter-Schwab) block of one pair record: MS:2912-2972. */
static void singular_pair_block(const hbo_geom *g, const int32_t *p, const duffy_rules *R, real k,
                                real kim, cx *slp, cx *dlp) {
    int test = p[0], trial = p[1], tm = p[2], sm = p[3], kind = p[4];
    int tl1 = p[5], tl2 = p[6], sl1 = p[7], sl2 = p[8];
    real nrm[3] = {tri_normal(g, trial, 0), tri_normal(g, trial, 1), tri_normal(g, trial, 2)};
    mirror3(nrm, sm);
    real jac = (2 * (real)g->areas[test]) * (2 * (real)g->areas[trial]);
    for (int i = 0; i < 3; ++i) slp[i] = cx_make(0, 0);
    for (int i = 0; i < 9; ++i) dlp[i] = cx_make(0, 0);
    int ki = kind - 1;
    for (int q = 0; q < R->n[ki]; ++q) {
        real txi, teta, sxi, seta, a, b;
        remap_singular(R->tx[ki][q], R->ty[ki][q], kind, tl1, tl2, &a, &b);
        image_ref_to_original(a, b, tm, &txi, &teta);
        remap_singular(R->sx[ki][q], R->sy[ki][q], kind, sl1, sl2, &a, &b);
        image_ref_to_original(a, b, sm, &sxi, &seta);
        real tp[3], sp[3];
        point_on_tri(g, test, txi, teta, tp);
        mirror3(tp, tm);
        point_on_tri(g, trial, sxi, seta, sp);
        mirror3(sp, sm);
        real tb[3] = {1 - txi - teta, txi, teta};
        real sb[3] = {1 - sxi - seta, sxi, seta};
        cx gg, dd;
        kern_g_dlp(sp[0] - tp[0], sp[1] - tp[1], sp[2] - tp[2], nrm[0], nrm[1], nrm[2], k, kim, &gg,
                   &dd);
        real w = R->w[ki][q] * jac;
        for (int i = 0; i < 3; ++i) {
            slp[i] = cx_add(slp[i], cx_scale(gg, tb[i] * w));
            for (int j = 0; j < 3; ++j)
                dlp[i * 3 + j] = cx_add(dlp[i * 3 + j], cx_scale(dd, tb[i] * sb[j] * w));
        }
    }
}

/* Exported pair block for tests.  mode 0 regular, 1 singular (pair record
 * built from the shared vertices; order = Duffy 1D order), 2 subdivided
 * (order = level).  out_slp[6] = 3 complex, out_dlp[18] = 3x3 complex. */
HBO_EXPORT int hbo_pair_block(const hbo_geom *g, int test, int trial, int tm, int sm, int mode,
                              int order, real k, real kim, real *out_slp, real *out_dlp) {
    cx slp[3], dlp[9];
    if (mode == 0) {
        regular_pair_block(g, test, trial, tm, sm, k, kim, slp, dlp);
    } else if (mode == 1) {
        shared_ids s = (g->symmetry == 0 && tm == 0 && sm == 0) ? shared_by_index(g, test, trial)
                                                                 : shared_by_key(g, test, trial, tm, sm);
        int64_t counts[3] = {0, 0, 0};
        int32_t rec[PAIR_W];
        if (append_pair(rec, 1, 0, test, trial, tm, sm, s, counts) != 1) return -1;
        duffy_rules R;
        rules_make(&R, order);
        singular_pair_block(g, rec, &R, k, kim, slp, dlp);
        rules_free(&R);
    } else if (mode == 2) {
        subdivided_pair_block(g, test, trial, tm, sm, k, kim, order, slp, dlp);
    } else {
        return -1;
    }
    for (int i = 0; i < 3; ++i) {
        out_slp[2 * i] = slp[i].re;
        out_slp[2 * i + 1] = slp[i].im;
    }
    for (int i = 0; i < 9; ++i) {
        out_dlp[2 * i] = dlp[i].re;
        out_dlp[2 * i + 1] = dlp[i].im;
    }
    return 0;
}

/* Apply delta = (block_new - regular) of every listed pair to A and rhs,
 * with delta accumulation in double per (row, col): MS:2974-3070
 * (applyDuffyCorrectionsCPU) for mode 1, MS:3072-3175
 * (applyNearFieldCorrectionsCPU) for mode 2.  Returns the number of unique
 * (row, col) slots touched. */
static int64_t apply_pair_deltas(const hbo_geom *g, const int32_t *pairs, int npairs, int mode,
                                 int param, real k, real kim, const real *q_re, const real *q_im,
                                 const real *beta_re, const real *beta_im, real *A_re, real *A_im,
                                 real *rhs_re, real *rhs_im) {
    int n = g->n_dofs;
    duffy_rules R;
    if (mode == 1) rules_make(&R, param);
    /* per pair: 3 slp deltas + 9 dlp deltas (complex) */
    double *delta = (double *)malloc(sizeof(double) * (size_t)npairs * 24);
#pragma omp parallel for schedule(dynamic, 16)
    for (int pi = 0; pi < npairs; ++pi) {
        const int32_t *p = pairs + (size_t)pi * PAIR_W;
        cx rs[3], rd[9], ns[3], nd[9];
        regular_pair_block(g, p[0], p[1], p[2], p[3], k, kim, rs, rd);
        if (mode == 1)
            singular_pair_block(g, p, &R, k, kim, ns, nd);
        else
            subdivided_pair_block(g, p[0], p[1], p[2], p[3], k, kim, param, ns, nd);
        double *o = delta + (size_t)pi * 24;
        for (int i = 0; i < 3; ++i) {
            cx d = cx_sub(ns[i], rs[i]);
            o[2 * i] = d.re;
            o[2 * i + 1] = d.im;
        }
        for (int i = 0; i < 9; ++i) {
            cx d = cx_sub(nd[i], rd[i]);
            o[6 + 2 * i] = d.re;
            o[6 + 2 * i + 1] = d.im;
        }
    }
    if (mode == 1) rules_free(&R);
    /* Dense FP64 scratch of the touched slots (accumulated in double as the
     * reference's triplet dictionary does, then added to A once). */
    double *acc = (double *)calloc((size_t)n * n * 2, sizeof(double));
    unsigned char *touched = (unsigned char *)calloc((size_t)n * n, 1);
    int64_t unique = 0;
    cx iK = cx_make(-kim, k);
    for (int pi = 0; pi < npairs; ++pi) {
        const int32_t *p = pairs + (size_t)pi * PAIR_W;
        const double *o = delta + (size_t)pi * 24;
        int test = p[0], trial = p[1];
        cx qt = cx_make(q_re ? q_re[trial] : 0, q_im ? q_im[trial] : 0);
        cx beta = cx_make(beta_re ? beta_re[trial] : 0, beta_im ? beta_im[trial] : 0);
        int has_robin = (beta.re != 0 || beta.im != 0);
        cx coupling = cx_mul(iK, beta);
        for (int i = 0; i < 3; ++i) {
            int row = p1_dof(g, test, i);
            cx sd = cx_make((real)o[2 * i], (real)o[2 * i + 1]);
            cx rd = cx_mul(sd, qt);
            rhs_re[row] += rd.re;
            rhs_im[row] += rd.im;
            for (int j = 0; j < 3; ++j) {
                int col = p1_dof(g, trial, j);
                cx dd = cx_make((real)o[6 + 2 * (i * 3 + j)], (real)o[6 + 2 * (i * 3 + j) + 1]);
                if (has_robin) dd = cx_sub(dd, cx_scale(cx_mul(sd, coupling), (real)(1.0 / 3.0)));
                size_t idx = (size_t)row * n + col;
                if (!touched[idx]) {
                    touched[idx] = 1;
                    unique++;
                }
                acc[2 * idx] += (double)dd.re;
                acc[2 * idx + 1] += (double)dd.im;
            }
        }
    }
    for (size_t idx = 0; idx < (size_t)n * n; ++idx)
        if (touched[idx]) {
            A_re[idx] += (real)acc[2 * idx];
            A_im[idx] += (real)acc[2 * idx + 1];
        }
    free(acc);
    free(touched);
    free(delta);
    return unique;
}

HBO_EXPORT int64_t hbo_apply_duffy(const hbo_geom *g, int order, real k, real kim, const real *q_re,
                                   const real *q_im, const real *beta_re, const real *beta_im,
                                   real *A_re, real *A_im, real *rhs_re, real *rhs_im,
                                   int64_t *counts) {
    int np = hbo_duffy_pairs(g, NULL, 0, counts);
    int32_t *pairs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1) * PAIR_W);
    hbo_duffy_pairs(g, pairs, np, counts);
    int64_t u = apply_pair_deltas(g, pairs, np, 1, order, k, kim, q_re, q_im, beta_re, beta_im, A_re,
                                  A_im, rhs_re, rhs_im);
    free(pairs);
    return u;
}

/* ---------------------------------------------------------------------- */
/* Near-field pair list: MS:2410-2558 (buildNearPairList)                  */
/* ---------------------------------------------------------------------- */

typedef struct {
    real c[3], lo[3], hi[3], edge;
} near_metric;

static near_metric tri_metric(const hbo_geom *g, int tri, int mask) {
    real p[3][3];
    for (int l = 0; l < 3; ++l) {
        int v = tri_vertex(g, tri, l);
        p[l][0] = vx(g, v, 0);
        p[l][1] = vx(g, v, 1);
        p[l][2] = vx(g, v, 2);
        mirror3(p[l], mask);
    }
    near_metric m;
    for (int c = 0; c < 3; ++c) {
        m.c[c] = (p[0][c] + p[1][c] + p[2][c]) / 3;
        real lo = p[0][c] < p[1][c] ? p[0][c] : p[1][c];
        m.lo[c] = lo < p[2][c] ? lo : p[2][c];
        real hi = p[0][c] > p[1][c] ? p[0][c] : p[1][c];
        m.hi[c] = hi > p[2][c] ? hi : p[2][c];
    }
    real e = 0;
    for (int l = 0; l < 3; ++l) {
        int o = (l + 1) % 3;
        real dx = p[l][0] - p[o][0], dy = p[l][1] - p[o][1], dz = p[l][2] - p[o][2];
        real d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > e) e = d2;
    }
    m.edge = RSQRT(e);
    return m;
}

static double axis_gap(real amin, real amax, real bmin, real bmax) {
    if (amax < bmin) return (double)(bmin - amax);
    if (bmax < amin) return (double)(amin - bmax);
    return 0.0;
}

HBO_EXPORT int hbo_near_pairs(const hbo_geom *g, double threshold, int32_t *out, int cap) {
    int T = g->n_triangles;
    int masks[4] = {0, 0, 0, 0};
    int nm = 1 + image_masks(g->symmetry, masks + 1);
    near_metric *M = (near_metric *)malloc(sizeof(near_metric) * (size_t)T * nm);
    for (int m = 0; m < nm; ++m)
        for (int t = 0; t < T; ++t) M[(size_t)m * T + t] = tri_metric(g, t, masks[m]);
    int n = 0;
    /* Ordered by (trial, trial mask, test mask, test) as MS:2548-2557. */
    for (int trial = 0; trial < T; ++trial)
        for (int smi = 0; smi < nm; ++smi)
            for (int tmi = 0; tmi < nm; ++tmi)
                for (int test = 0; test < T; ++test) {
                    near_metric *a = &M[(size_t)tmi * T + test], *b = &M[(size_t)smi * T + trial];
                    double cut = threshold * (double)(a->edge > b->edge ? a->edge : b->edge);
                    double cut2 = cut * cut;
                    double dx = axis_gap(a->lo[0], a->hi[0], b->lo[0], b->hi[0]);
                    double dy = axis_gap(a->lo[1], a->hi[1], b->lo[1], b->hi[1]);
                    double dz = axis_gap(a->lo[2], a->hi[2], b->lo[2], b->hi[2]);
                    if (dx * dx + dy * dy + dz * dz > cut2) continue;
                    double cx0 = (double)(a->c[0] - b->c[0]), cy0 = (double)(a->c[1] - b->c[1]),
                           cz0 = (double)(a->c[2] - b->c[2]);
                    if (cx0 * cx0 + cy0 * cy0 + cz0 * cz0 >= cut2) continue;
                    shared_ids s = shared_by_key(g, test, trial, masks[tmi], masks[smi]);
                    if (s.count > 0) continue;
                    if (out && n < cap) {
                        int32_t *p = out + (size_t)n * PAIR_W;
                        p[0] = test; p[1] = trial; p[2] = masks[tmi]; p[3] = masks[smi];
                        p[4] = 0; p[5] = p[6] = p[7] = p[8] = 0;
                    }
                    n++;
                }
    free(M);
    return n;
}

HBO_EXPORT int64_t hbo_apply_near(const hbo_geom *g, int level, double threshold, real k, real kim,
                                  const real *q_re, const real *q_im, const real *beta_re,
                                  const real *beta_im, real *A_re, real *A_im, real *rhs_re,
                                  real *rhs_im) {
    int np = hbo_near_pairs(g, threshold, NULL, 0);
    int32_t *pairs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1) * PAIR_W);
    hbo_near_pairs(g, threshold, pairs, np);
    apply_pair_deltas(g, pairs, np, 2, level, k, kim, q_re, q_im, beta_re, beta_im, A_re, A_im,
                      rhs_re, rhs_im);
    free(pairs);
    return np;
}

/* ---------------------------------------------------------------------- */
/* Exterior field: MS:1843-1895 (evaluateExteriorReference)                */
/* ---------------------------------------------------------------------- */

HBO_EXPORT void hbo_field(const hbo_geom *g, real k, const real *p_re, const real *p_im,
                          const real *q_re, const real *q_im, int n_obs, const float *obs3xn,
                          real *out_re, real *out_im) {
    int T = g->n_triangles;
    int masks[3];
    int nm = image_masks(g->symmetry, masks);
    real qx[6], qy[6], qw[6];
    hbo_rule6(qx, qy, qw);
#pragma omp parallel for schedule(dynamic, 1)
    for (int o = 0; o < n_obs; ++o) {
        real ox = (real)obs3xn[o], oy = (real)obs3xn[n_obs + o], oz = (real)obs3xn[2 * n_obs + o];
        cx acc = cx_make(0, 0);
        for (int tri = 0; tri < T; ++tri) {
            real nrm[3] = {tri_normal(g, tri, 0), tri_normal(g, tri, 1), tri_normal(g, tri, 2)};
            real jac = 2 * (real)g->areas[tri];
            cx qt = cx_make(q_re[tri], q_im[tri]);
            for (int a = 0; a < 6; ++a) {
                real sp[3];
                point_on_tri(g, tri, qx[a], qy[a], sp);
                real b[3] = {1 - qx[a] - qy[a], qx[a], qy[a]};
                cx ps = cx_make(0, 0);
                for (int l = 0; l < 3; ++l) {
                    int d = p1_dof(g, tri, l);
                    ps = cx_add(ps, cx_scale(cx_make(p_re[d], p_im[d]), b[l]));
                }
                cx gg, dd;
                kern_g_dlp(sp[0] - ox, sp[1] - oy, sp[2] - oz, nrm[0], nrm[1], nrm[2], k, 0, &gg, &dd);
                acc = cx_add(acc, cx_scale(cx_sub(cx_mul(dd, ps), cx_mul(gg, qt)), qw[a] * jac));
                for (int m = 0; m < nm; ++m) {
                    real ip[3] = {sp[0], sp[1], sp[2]};
                    real in[3] = {nrm[0], nrm[1], nrm[2]};
                    mirror3(ip, masks[m]);
                    mirror3(in, masks[m]);
                    kern_g_dlp(ip[0] - ox, ip[1] - oy, ip[2] - oz, in[0], in[1], in[2], k, 0, &gg,
                               &dd);
                    acc = cx_add(acc,
                                 cx_scale(cx_sub(cx_mul(dd, ps), cx_mul(gg, qt)), qw[a] * jac));
                }
            }
        }
        out_re[o] = acc.re;
        out_im[o] = acc.im;
    }
}

/* CHIEF rows: MS:1930-2036 (assembleChiefRows).  For interior point i and
 * P1 column j: C[i][j] = sum over triangles and images of the 6-point
 * integral of dG/dn_y (real k) times the basis, minus the Robin fold
 * (i k beta / 3) int G; d[s][i] = sum_t int G * q_s[t].  C is row-major m x n
 * complex, d is [n_src][m] complex. */
HBO_EXPORT void hbo_chief_rows(const hbo_geom *g, int m, const float *pts3xm, real k, real kim,
                               int n_src, const real *q_re, const real *q_im, const real *beta_re,
                               const real *beta_im, real *C_re, real *C_im, real *d_re,
                               real *d_im) {
    int T = g->n_triangles, n = g->n_dofs;
    int masks[3];
    int nm = image_masks(g->symmetry, masks);
    real qx[6], qy[6], qw[6];
    hbo_rule6(qx, qy, qw);
    cx iK = cx_make(-kim, k);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ci = 0; ci < m; ++ci) {
        real ox = (real)pts3xm[ci], oy = (real)pts3xm[m + ci], oz = (real)pts3xm[2 * m + ci];
        for (int tri = 0; tri < T; ++tri) {
            real nrm[3] = {tri_normal(g, tri, 0), tri_normal(g, tri, 1), tri_normal(g, tri, 2)};
            real jac = 2 * (real)g->areas[tri];
            cx dcoef[3] = {cx_make(0, 0), cx_make(0, 0), cx_make(0, 0)};
            cx gsum = cx_make(0, 0);
            for (int a = 0; a < 6; ++a) {
                real sp[3];
                point_on_tri(g, tri, qx[a], qy[a], sp);
                real b[3] = {1 - qx[a] - qy[a], qx[a], qy[a]};
                real w = qw[a] * jac;
                for (int mi = -1; mi < nm; ++mi) {
                    real ip[3] = {sp[0], sp[1], sp[2]};
                    real in[3] = {nrm[0], nrm[1], nrm[2]};
                    if (mi >= 0) {
                        mirror3(ip, masks[mi]);
                        mirror3(in, masks[mi]);
                    }
                    cx gg, dd;
                    kern_g_dlp(ip[0] - ox, ip[1] - oy, ip[2] - oz, in[0], in[1], in[2], k, 0, &gg,
                               &dd);
                    for (int l = 0; l < 3; ++l) dcoef[l] = cx_add(dcoef[l], cx_scale(dd, b[l] * w));
                    gsum = cx_add(gsum, cx_scale(gg, w));
                }
            }
            for (int s = 0; s < n_src; ++s) {
                cx q = cx_make(q_re[(size_t)s * T + tri], q_im[(size_t)s * T + tri]);
                cx v = cx_mul(gsum, q);
                d_re[(size_t)s * m + ci] += v.re;
                d_im[(size_t)s * m + ci] += v.im;
            }
            cx beta = cx_make(beta_re ? beta_re[tri] : 0, beta_im ? beta_im[tri] : 0);
            int has_robin = beta.re != 0 || beta.im != 0;
            cx fold = cx_scale(cx_mul(gsum, cx_mul(iK, beta)), (real)(1.0 / 3.0));
            for (int l = 0; l < 3; ++l) {
                int col = p1_dof(g, tri, l);
                cx term = has_robin ? cx_sub(dcoef[l], fold) : dcoef[l];
                C_re[(size_t)ci * n + col] += term.re;
                C_im[(size_t)ci * n + col] += term.im;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Surface reductions: MS:807-841 (averageSurfacePressureForTag)           */
/* ---------------------------------------------------------------------- */

HBO_EXPORT void hbo_surface_average(const hbo_geom *g, const int32_t *tags, int tag,
                                    const real *p_re, const real *p_im, double *out3) {
    double sr = 0, si = 0, ta = 0;
    for (int tri = 0; tri < g->n_triangles; ++tri) {
        if (tags[tri] != tag) continue;
        double area = (double)g->areas[tri];
        if (area <= 0) continue;
        int d0 = p1_dof(g, tri, 0), d1 = p1_dof(g, tri, 1), d2 = p1_dof(g, tri, 2);
        double ar = ((double)p_re[d0] + (double)p_re[d1] + (double)p_re[d2]) / 3.0;
        double ai = ((double)p_im[d0] + (double)p_im[d1] + (double)p_im[d2]) / 3.0;
        sr += ar * area;
        si += ai * area;
        ta += area;
    }
    if (ta < 1e-30) {
        out3[0] = out3[1] = out3[2] = 0;
        return;
    }
    out3[0] = sr / ta;
    out3[1] = si / ta;
    out3[2] = ta;
}

HBO_EXPORT int hbo_real_bytes(void) { return (int)sizeof(real); }

HBO_EXPORT int hbo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Contributions of every (test, trial, masks) pair that has its test or its
 * trial triangle in sel (regular rule, mass term, Sauter-Schwab deltas),
 * added to A and rhs: the part of the corrected system that moves with the
 * vertices of the selected triangles (local finite differences).  Returns
 * the number of touching pairs corrected. */
HBO_EXPORT int64_t hbo_assemble_subset(const hbo_geom *g, const unsigned char *sel, int order,
                                       real k, real kim, const real *q_re, const real *q_im,
                                       const real *beta_re, const real *beta_im, real *A_re,
                                       real *A_im, real *rhs_re, real *rhs_im) {
    assemble_regular(g, sel, k, kim, q_re, q_im, beta_re, beta_im, A_re, A_im, rhs_re, rhs_im);
    int64_t counts[3] = {0, 0, 0};
    int np = hbo_duffy_pairs(g, NULL, 0, counts);
    int32_t *pairs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1) * PAIR_W);
    hbo_duffy_pairs(g, pairs, np, counts);
    int nk = 0;
    for (int i = 0; i < np; ++i) {
        const int32_t *p = pairs + (size_t)i * PAIR_W;
        if (sel[p[0]] || sel[p[1]]) {
            if (nk != i) memmove(pairs + (size_t)nk * PAIR_W, p, sizeof(int32_t) * PAIR_W);
            nk++;
        }
    }
    apply_pair_deltas(g, pairs, nk, 1, order, k, kim, q_re, q_im, beta_re, beta_im, A_re, A_im,
                      rhs_re, rhs_im);
    free(pairs);
    return nk;
}
