// Internal declarations shared by the host plan builder, the kernels and the
// C ABI of libhbem.  Nothing here is part of the public ABI (include/hbem.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hbem {

// ---------------------------------------------------------------------------
// Quadrature constants shared by host and device.
// ---------------------------------------------------------------------------

// Degree-4 six-point triangle rule (weights sum to 1/2), the rule of the
// reference's triangleRule6 (main.swift:1472-1489).
constexpr double kRule6X[6] = {0.4459484909159651, 0.0915762135097710, 0.1081030181680700,
                               0.4459484909159651, 0.8168475729804590, 0.0915762135097710};
constexpr double kRule6Y[6] = {0.4459484909159651, 0.0915762135097700, 0.4459484909159651,
                               0.1081030181680700, 0.0915762135097700, 0.8168475729804580};
constexpr double kRule6W[6] = {0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220,
                               0.5 * 0.2233815896780110, 0.5 * 0.2233815896780110,
                               0.5 * 0.1099517436553220, 0.5 * 0.1099517436553220};

constexpr int kMaxMasks = 4;  // identity + up to 3 image masks

// Per-triangle records.  The regular-assembly kernel reads the trial
// triangle per lane (structure-of-arrays, coalesced) and the test triangle
// per wave (array-of-structures, scalar loads).  Quadrature points are
// evaluated once on the host in double precision and rounded to float.
//
// Float record (kTriF floats):
//   [0..17]  quadrature points x_a (a = 0..5), xyz
//   [18..20] unit normal
//   [21]     area
//   [22]     2 * area (Jacobian of the reference map)
//   [23..31] vertices v0, v1, v2 (xyz each)
constexpr int kTriF = 32;
// Int record (kTriI ints):
//   [0..2]   P1 dofs
//   [3..5]   vertex indices
//   [6+3m .. 8+3m] coordinate-key ids of the vertices under mask index m
//   [18]     physical tag
//   [19]     original (caller) triangle index
//   [20..22] position of each dof in the triangle's tile dof list
//   [23]     pad
constexpr int kTriI = 24;

// Triangles are stored in Morton order of their centroids and grouped into
// tiles of up to kTileTris consecutive triangles with at most kTileMaxDofs
// distinct dofs; the regular-assembly kernel accumulates one tile x tile
// block of the matrix in LDS.
constexpr int kTileTris = 64;
constexpr int kTileMaxDofs = 90;  // 90^2 complex floats = 64800 B of LDS

struct Plan {
    int n_vertices = 0;
    int n_triangles = 0;
    int n_dofs = 0;
    int symmetry = 0;
    int n_masks = 1;                 // 1 + number of image masks
    int masks[kMaxMasks] = {0, 0, 0, 0};
    float row_weight = 1.0f;         // 2^planes

    std::vector<float> tri_f_aos;    // [T][kTriF]
    std::vector<int32_t> tri_i_aos;  // [T][kTriI]
    std::vector<float> tri_f_soa;    // [kTriF][T]
    std::vector<int32_t> tri_i_soa;  // [kTriI][T]

    // Singular pairs sorted by kind (1, 2, 3).  Each pair is two ints:
    //   pair_tt[2p] = test, pair_tt[2p+1] = trial
    //   pair_code[p] = tm | sm<<3 | kind<<6 | tl1<<8 | tl2<<10 | sl1<<12 | sl2<<14
    std::vector<int32_t> pair_tt;
    std::vector<int32_t> pair_code;
    int64_t kind_count[3] = {0, 0, 0};
    int64_t image_pairs = 0;
    int64_t unique_slots = 0;

    // Sauter-Schwab rule tables for kinds 1..3, concatenated: per point
    // (test x, test y, trial x, trial y, weight).
    std::vector<float> rule;         // [n_points][5]
    int rule_offset[3] = {0, 0, 0};
    int rule_count[3] = {0, 0, 0};

    std::vector<float> areas;        // [T]
    std::vector<int32_t> tags;       // [T]
    std::vector<int32_t> l2g;        // [T][3]

    // Tiles of device-order triangles.
    std::vector<int32_t> tile_start;   // [n_tiles]
    std::vector<int32_t> tile_count;   // [n_tiles]
    std::vector<int32_t> tile_ndofs;   // [n_tiles]
    std::vector<int32_t> tile_dofs;    // [n_tiles][kTileMaxDofs], -1 padded
    int max_tile_dofs = 0;
};

// Builds the plan.  Returns an empty string on success, else a message.
std::string build_plan(int n_vertices, int n_triangles, int n_dofs, int symmetry,
                       const float *vertices_3xn, const int32_t *triangles_3xm,
                       const int32_t *l2g, const float *areas, const float *normals_3xm,
                       const int32_t *tags, Plan *plan);

// Sauter-Schwab tables of one kind (1D Gauss order 4).  Returns the number of
// points written (rows of 5 floats).
int duffy_rule_f32(int kind, std::vector<float> *out);

// ---------------------------------------------------------------------------
// Device-side launch wrappers (each enqueues on `stream`, no sync).
// ---------------------------------------------------------------------------

struct DeviceGeometry {
    const float *tri_f_aos;
    const int32_t *tri_i_aos;
    const float *tri_f_soa;
    const int32_t *tri_i_soa;
    const int32_t *pair_tt;
    const int32_t *pair_code;
    const float *rule;
    int rule_offset[3];
    int rule_count[3];
    int n_pairs;
    int n_triangles;
    int n_dofs;
    int symmetry;
    int n_masks;
    int masks[kMaxMasks];
    float row_weight;
    const int32_t *tile_start;
    const int32_t *tile_count;
    const int32_t *tile_ndofs;
    const int32_t *tile_dofs;
    int n_tiles;
    int max_tile_dofs;
};

// Zero n floats.
void launch_fill_zero(float *ptr, size_t n, hipStream_t stream);
void launch_fill_zero_f64(double *ptr, size_t n, hipStream_t stream);

// Regular (non-touching) pair quadrature of all pairs, all frequencies of a
// batch, accumulated into A (column-major, interleaved complex, lda) and
// rhs (interleaved complex, stride ldr per frequency).
void launch_assemble_regular(const DeviceGeometry &g, int n_batch, const float *k_real,
                             const float *k_imag, const float *q /* [B][T][2] */, float *A,
                             int64_t lda, int64_t a_stride, float *rhs, int64_t rhs_stride,
                             hipStream_t stream);

// Singular (Sauter-Schwab) pair blocks plus the -M/2 mass term.
void launch_assemble_singular(const DeviceGeometry &g, int n_batch, const float *k_real,
                              const float *k_imag, const float *q, float *A, int64_t lda,
                              int64_t a_stride, float *rhs, int64_t rhs_stride,
                              hipStream_t stream);

// Exterior field p(x_o) for each frequency of a batch.
void launch_field(const DeviceGeometry &g, int n_batch, const float *field_k,
                  const float *p /* [B][p_stride] complex */, int64_t p_stride,
                  const float *q /* [B][T] complex */, int n_obs, const float *obs_3xn,
                  float *out /* [B][n_obs] complex */, hipStream_t stream);

// Area-weighted tag averages (FP64): out[b][t] = (re, im) for tags[t].
void launch_tag_average(const DeviceGeometry &g, int n_batch, const float *p, int64_t p_stride,
                        int n_tags, const int32_t *tags_dev, double *out, hipStream_t stream);

// ---------------------------------------------------------------------------
// Batched complex LU (column-major interleaved complex, padded order N).
// ---------------------------------------------------------------------------

struct LuWorkspace {
    float *A;          // [B][lda*N] complex (factored in place)
    int64_t lda;
    int64_t a_stride;  // complex elements between matrices
    int N;             // padded order (multiple of kLuPad)
    int n;             // true order
    int B;
    int32_t *ipiv;     // [B][N]
    int32_t *info;     // [B]
};

constexpr int kLuPad = 64;

// Pads rows/cols n..N-1 with the identity (zeros elsewhere).
void launch_pad_identity(float *A, int64_t lda, int64_t a_stride, int n, int N, int B,
                         hipStream_t stream);

// 1-norm of each matrix (max column sum of |a_ij|) -> norms[B] (double).
void launch_norm1(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *norms,
                  hipStream_t stream);

void lu_factor(const LuWorkspace &w, hipStream_t stream);

// Solves A x = b (trans = 0) or A^H x = b (trans = 2) in place on x
// [B][ldx] interleaved complex, using the LU factors and pivots.
void lu_solve(const LuWorkspace &w, float *x, int64_t ldx, int trans, hipStream_t stream);

// Residual r = b - A x in FP64 with A the FP32 operator (column-major
// interleaved), x and r FP64 interleaved.  Also writes per-system
// ||r||_inf and ||b||_inf into norms[2*b], norms[2*b+1].
void launch_residual_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                         const float *b, int64_t ldb, const double *x, int64_t ldx, double *r,
                         double *norms, hipStream_t stream);

// Helpers for refinement and rcond estimation.
void launch_copy_f32_to_f64(const float *src, double *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream);
void launch_copy_f64_to_f32(const double *src, float *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream);
void launch_axpy_f32_to_f64(const float *d, double *x, int64_t ld_d, int64_t ld_x, int n, int B,
                            const int32_t *active, hipStream_t stream);

// Copy a row-major split (re, im) host-layout matrix block on device into
// the column-major interleaved LU layout and back.
void launch_rowmajor_split_to_colmajor(const float *re, const float *im, int n, int B, float *A,
                                       int64_t lda, int64_t a_stride, hipStream_t stream);
void launch_colmajor_to_rowmajor_split(const float *A, int64_t lda, int64_t a_stride, int n,
                                       int B, float *re, float *im, hipStream_t stream);

}  // namespace hbem
