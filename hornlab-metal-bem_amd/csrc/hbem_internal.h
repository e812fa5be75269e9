// Internal declarations shared by the host plan builder, the kernels and the
// C ABI of libhbem.  Nothing here is part of the public ABI (include/hbem.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace hbem {

// The C ABI's error convention for entry points defined outside hbem_api.cpp:
// run_guarded returns HBEM_OK or the code of the failure and records its
// message for hbem_last_error(); fail_with / check_hip raise such failures.
int32_t run_guarded(const std::function<void()> &f);
[[noreturn]] void fail_with(int code, const std::string &msg);
void check_hip(hipError_t e, const char *what);

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
//   [23]     device index (test-side records), else -1
constexpr int kTriI = 24;

// Triangles are stored in Morton order of their centroids and grouped into
// tiles of up to kTileTris consecutive triangles with at most kTileMaxDofs
// distinct dofs; the regular-assembly kernel accumulates the two tile x
// tile blocks of a tile pair in LDS (2 * 90^2 complex floats = 129600 B at
// most).
constexpr int kTileTris = 64;
constexpr int kTileMaxDofs = 90;

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
    // Test-side records of the regular kernel ([T][kTriF], [T][kTriI]):
    // position t0 + 16 w + i of a tile holds the i-th test of wave w, the
    // wave's tests ordered as a path through shared edges, each record's
    // vertices permuted so that a dof shared with the previous test keeps
    // its local position; int word 23 is the test's device index.
    std::vector<float> tri_t_f;
    std::vector<int32_t> tri_t_i;

    // Counts of the full singular pair list (every touching (test image,
    // trial image) combination), as the reference reports them.
    int64_t kind_count[3] = {0, 0, 0};
    int64_t image_pairs = 0;
    int64_t unique_slots = 0;

    // Singular pair classes.  Two touching combinations (tm, sm) and
    // (tm', sm') of one (test, trial) with tm^sm == tm'^sm', the same mask
    // parity and the same shared locals differ by one isometry applied to
    // both triangles, so their blocks are equal; each class is evaluated
    // once with its size as weight.  Classes are sorted by kind and every
    // kind is padded to a multiple of 64 (one lane per class, kind-uniform
    // waves).  For class c the quadrature point of rule coordinates
    // (x, y | x', y') has
    //   r = Q - P = R0 + (x'-x) E1 + (y'-y) E2 + x' G1 + y' G2
    // with P(x,y) = P0 + x E1 + y E2 the test image map (shared-vertex remap
    // and odd-mask swap folded in), Q = Q0 + x' F1 + y' F2 the trial image
    // map, R0 = Q0 - P0, G1 = F1 - E1, G2 = F2 - E2, all formed in double;
    // R0, G1, G2 vanish on shared vertices and edges.  Rule-local basis a of
    // the test (b of the trial) is the P1 function of the original local
    // vertex that the map sends rule vertex a to.
    // Float fields, [kSgF][sg_np]:
    //   [0..2] R0  [3..5] E1  [6..8] E2  [9..11] G1  [12..14] G2
    //   [15..17] trial image normal  [18] (2 A_t)(2 A_s) * weight
    //   [19] mass coefficient -A_t/2 * weight on coincident self pairs, else 0
    // Int fields, [kSgI][sg_np]:
    //   [0..2] row dof of test basis a  [3..5] column dof of trial basis b
    //   [6] trial original index (-1 on padding)
    std::vector<float> sg_f;
    std::vector<int32_t> sg_i;
    // Adjoint fields [kSgAdj][sg_np]: [0..2] test vertex (caller index) of
    // the remapped test corners, [3..5] the same for the trial, [6] test mask
    // | trial mask << 4, [7] test caller index.
    std::vector<int32_t> sg_adj;
    int sg_np = 0;                 // padded class count
    int sg_waves[3] = {0, 0, 0};   // 64-lane groups per kind
    int64_t sg_classes = 0;        // unpadded class count

    // Sauter-Schwab rules of kinds 1..3, concatenated; per point kSgRule
    // floats: (x'-x, y'-y, x', y', w rho_a rho'_b (a, b = 0..2, 9 values),
    // w rho_a (3 values)) with rho = (1-x-y, x, y), rho' = (1-x'-y', x', y').
    std::vector<float> sg_rule;
    int sg_rule_offset[3] = {0, 0, 0};  // in points
    int sg_rule_count[3] = {0, 0, 0};
    // The same classes under the plain 6 x 6 rule (regular quadrature on the
    // touching pairs too, the reference's "optimized" assembly): per kind the
    // 36 tensor points in the layout above, padded with zero-weight points to
    // a whole number of FP32 runs.  The 6-point rule maps onto itself under
    // every vertex permutation, so the class maps leave it unchanged.
    std::vector<float> sg_rule_reg;
    int sg_rule_reg_offset[3] = {0, 0, 0};
    int sg_rule_reg_count[3] = {0, 0, 0};

    std::vector<float> areas;        // [T]
    std::vector<int32_t> tags;       // [T]
    std::vector<int32_t> l2g;        // [T][3]
    std::vector<float> vtx;          // [3][nv] caller vertices
    std::vector<int32_t> tri;        // [3][T] caller triangles
    std::vector<float> normals;      // [3][T]

    // Near-field classes (built on demand for a threshold and a level):
    // non-touching (test, trial image) combinations whose mirrored triangles
    // lie within threshold x the longer longest edge (bounding-box gap and
    // centroid distance), grouped like the singular classes by
    // (test, trial, tm^sm) with the count as weight.  Float fields
    // [kNcF][nc_np]: [0..2] R0 = Q0 - P0, [3..5] E1, [6..8] E2 (test, identity),
    // [9..11] F1, [12..14] F2 (trial image), [15..17] trial image normal,
    // [18] (2 A_t)(2 A_s) weight.  Int fields [kNcI][nc_np]: [0..2] row dofs,
    // [3..5] column dofs, [6] trial caller index (-1 padding).  Rule rows
    // (x, y, w): the level-L subdivided 6-point rule, then the 6-point rule.
    int nc_level = 0;
    double nc_threshold = 0.0;
    std::vector<float> nc_f;
    std::vector<int32_t> nc_i;
    int nc_np = 0;
    int64_t nc_pairs = 0;     // near (test image, trial image) combinations
    std::vector<float> nc_rule;
    int nc_sub_points = 0;

    // Tiles of device-order triangles.
    std::vector<int32_t> tile_start;   // [n_tiles]
    std::vector<int32_t> tile_count;   // [n_tiles]
    std::vector<int32_t> tile_ndofs;   // [n_tiles]
    std::vector<int32_t> tile_dofs;    // [n_tiles][kTileMaxDofs], -1 padded
    std::vector<int32_t> tile_pairs;   // [n_tile_pairs][2] = (S, T), S >= T
    int max_tile_dofs = 0;
};

// Builds the plan.  Returns an empty string on success, else a message.
std::string build_plan(int n_vertices, int n_triangles, int n_dofs, int symmetry,
                       const float *vertices_3xn, const int32_t *triangles_3xm,
                       const int32_t *l2g, const float *areas, const float *normals_3xm,
                       const int32_t *tags, Plan *plan);

constexpr int kNcF = 19;
constexpr int kNcI = 7;

// Builds the near-field classes of Plan for (threshold, level); returns the
// number of classes.
int64_t build_near_classes(Plan *P, double threshold, int level);

constexpr int kSgAdj = 8;
constexpr int kSgF = 20;
constexpr int kSgI = 7;
constexpr int kSgRule = 16;

// Sauter-Schwab points of one kind (1D Gauss order 4) in double: rows of
// (x, y, x', y', w).  Returns the number of points written.
int duffy_rule(int kind, std::vector<double> *out);

// ---------------------------------------------------------------------------
// Device-side launch wrappers (each enqueues on `stream`, no sync).
// ---------------------------------------------------------------------------

struct DeviceGeometry {
    const float *tri_f_aos;
    const int32_t *tri_i_aos;
    const float *tri_t_f;
    const int32_t *tri_t_i;
    const float *tri_f_soa;
    const int32_t *tri_i_soa;
    const float *sg_f;
    const int32_t *sg_i;
    const float *sg_rule;
    int sg_np;
    int sg_waves[3];
    int sg_rule_offset[3];
    int sg_rule_count[3];
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
    const int32_t *tile_pairs;
    const int32_t *l2g;  // [T][3] P1 dofs in caller triangle order
    const int32_t *sg_adj;
    const float *vtx;      // [3][nv] caller vertices
    const int32_t *tri_v;  // [3][T] caller triangles
    const float *normals;  // [3][T] caller normals
    const float *jac_c;    // [T] 2 * area, caller order
    int n_vertices;
    const float *nc_f;
    const int32_t *nc_i;
    const float *nc_rule;
    int nc_np;
    int nc_sub_points;
    int n_tiles;
    int n_tile_pairs;
    int max_tile_dofs;
};

// Transcendental-pipe probe: blocks x 256 threads x iters x 16 ops
// (sin, cos, rsq in equal parts), out [blocks * 256].
void launch_probe_transcendental(float *out, int blocks, int iters, hipStream_t stream);

// Writes the device's constant-rate wall clock (wall_clock64) to *slot when
// the stream reaches it (timestamps inside captured graphs).
void launch_timestamp(long long *slot, hipStream_t stream);

// Zero n floats.
void launch_fill_zero(float *ptr, size_t n, hipStream_t stream);
void launch_fill_zero_f64(double *ptr, size_t n, hipStream_t stream);

// Regular (non-touching) pair quadrature of all pairs, all frequencies of a
// batch, accumulated into A (column-major, interleaved complex, lda) and
// rhs (interleaved complex, stride rhs_stride per vector, n_src vectors per
// frequency).  q [B][n_src][T] complex; beta [B][T] complex Robin admittance
// or null.
void launch_assemble_regular(const DeviceGeometry &g, int n_batch, const float *k_real,
                             const float *k_imag, const float *q, int n_src, const float *beta,
                             float *A, int64_t lda, int64_t a_stride, float *rhs,
                             int64_t rhs_stride, hipStream_t stream, bool precise = false);

// One member of a multi-mesh assembly launch (a session group solved
// together): its geometry, its batch's inputs, where its systems go, and
// the first flattened block it owns in the regular and singular launches.
struct AssemblyJob {
    DeviceGeometry g;
    const float *k_real;
    const float *k_imag;  // null: real k
    const float *q;       // [n_batch][n_src][T] complex
    const float *beta;    // null: no Robin walls
    int n_src;
    int n_batch;
    float *A;
    int64_t lda, a_stride;
    float *rhs;
    int64_t rhs_stride;
    int block_begin_reg;
    int block_begin_sing;
};

// Blocks a member needs in the multi-mesh launches.
inline int regular_blocks(const DeviceGeometry &g, int n_batch) { return g.n_tile_pairs * n_batch; }
inline int singular_blocks(const DeviceGeometry &g, int n_batch) {
    const int waves = g.sg_waves[0] + g.sg_waves[1] + g.sg_waves[2];
    return ((waves + 3) / 4) * ((n_batch + 1) / 2);
}

// Regular and singular assembly of every job in one launch each (jobs_dev:
// the jobs in device memory, jobs: the same on the host, with block_begin_*
// prefix sums of regular_blocks / singular_blocks).  All jobs share
// n_masks; complex_k selects the damped kernels for all of them.
void launch_assemble_multi(const AssemblyJob *jobs_dev, const std::vector<AssemblyJob> &jobs,
                           bool complex_k, hipStream_t stream);

// Singular (Sauter-Schwab) pair blocks plus the -M/2 mass term.
void launch_assemble_singular(const DeviceGeometry &g, int n_batch, const float *k_real,
                              const float *k_imag, const float *q, int n_src, const float *beta,
                              float *A, int64_t lda, int64_t a_stride, float *rhs,
                              int64_t rhs_stride, hipStream_t stream);

// Near-field deltas (subdivided minus 6x6 rule) of the near classes.
void launch_assemble_near(const DeviceGeometry &g, int n_batch, const float *k_real,
                          const float *k_imag, const float *q, int n_src, const float *beta,
                          float *A, int64_t lda, int64_t a_stride, float *rhs,
                          int64_t rhs_stride, hipStream_t stream);

// Adjoint (shape derivative) stages, all frequencies of a batch per launch.
// Cotangent grad_field [B][n_obs] complex float; c [B][c_stride] and ft
// [B][T] complex double receive E^H g and F^H g before conjugation; gV
// [3][nv] and gN [3][T] (double) the vertex and normal gradients.
void launch_field_adjoint(const DeviceGeometry &g, int n_batch, const float *field_k,
                          const float *p, int64_t p_stride, const float *qt, int n_obs,
                          const float *obs, const float *grad_field, double *c, int64_t c_stride,
                          double *ft, double *gV, double *gN, hipStream_t stream);
// e = conj(c) (+ the Robin term from ft), as float and double, zero-padded
// to N per system.
void launch_adjoint_rhs(const DeviceGeometry &g, int n_batch, const float *k_real,
                        const float *k_imag, const float *beta, const double *ft, double *c,
                        int64_t ld, int n, int N, float *e32, double *e64, hipStream_t stream);
// r = e - A^H x in FP64 for B vectors (vector v uses matrix v / vpm),
// norms[2 v] = max|r|, norms[2 v + 1] = max|e|.
void launch_residual_h_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                           const double *e, const double *x, int64_t ldx, double *r,
                           double *norms, hipStream_t stream, int vpm = 1);

// CHIEF rows C [B][m][n] and d [B][n_src][m] (interleaved complex float,
// accumulated: zero them first) of m interior points pts [3][m].
void launch_chief_rows(const DeviceGeometry &g, int n_batch, int m, const float *pts,
                       const float *k_real, const float *k_imag, int n_src, const float *q,
                       const float *beta, float *C, float *d, hipStream_t stream);
// ||A||_inf per matrix (FP64).
void launch_row_abs_sum(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *out,
                        hipStream_t stream);
// Regular and singular pair derivatives of Re(lambda^H (V qt - K p)).
void launch_assembly_adjoint(const DeviceGeometry &g, int n_batch, const float *k_real,
                             const float *k_imag, const float *lam, const float *p, int64_t ld,
                             const float *qt, double *gV, double *gN, double *sig,
                             hipStream_t stream);
// Mass term, normal -> vertex gradients, grad_q = conj(ft + sig).
void launch_adjoint_finish(const DeviceGeometry &g, int n_batch, const float *lam, const float *p,
                           int64_t ld, const double *gN, double *gV, const double *ft,
                           const double *sig, float *grad_q, hipStream_t stream);

// qt = q + i k beta mean(p) per triangle for n_vec vectors (vpk per
// frequency): the Neumann data of the field with Robin walls.
void launch_robin_neumann(const DeviceGeometry &g, int n_vec, int vpk, const float *k_real,
                          const float *k_imag, const float *q, const float *beta, const float *p,
                          int64_t p_stride, float *qt, hipStream_t stream);

// Exterior field p(x_o) for n_batch vectors (vector v at wavenumber
// field_k[v / vpk]).
void launch_field(const DeviceGeometry &g, int n_batch, const float *field_k, int vpk,
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
    int32_t *gl;       // [B][N / 32][kLuGather] composed row exchanges per step
    int32_t *scratch_i;  // [B][kLuMaxOrder / 32] tournament candidates
    float *scratch_f;    // [B][32][32] + [B][64][32] complex: top block, gathered rows
    uint16_t *pack = nullptr;  // [B][lu_pack_u16(N)]: 16-bit trailing operands (or null)
    // Row maps, final-order L and diagonal-block inverses of the
    // multi-workgroup solves (or null; then lu_factor skips them): perm[B][N]
    // original row of each final row, pi[B][lu_solve_blocks(N)][N] row of
    // each final row in the order at the end of outer block s, lf[B][a_stride]
    // complex L below the diagonal in final row order (layout of A),
    // dinv[B][N / 64][2][64][64] complex inverses of the 64 x 64 diagonal
    // blocks of L and U.
    int32_t *perm = nullptr;
    int32_t *pi = nullptr;
    float *lf = nullptr;
    float *dinv = nullptr;
};

constexpr int kLuPad = 64;
constexpr int kLuOuter = 256;   // outer block of the two-level LU (exchange granularity)
constexpr int kLuGather = 132;  // ints per step of LuWorkspace::gl
constexpr int kLuMaxOrder = 32768;  // tournament panels nominate from <= 32 chunks
// orders up to this solve with one 1024-thread block per vector (x in LDS);
// larger ones in per-outer-block steps
constexpr int kLuSolveOneBlock = 2048;

// orders up to this solve on N / 64 workgroups per vector when the caller
// gives a TrsvSync (the row-map kernel keeps 2 N ints in LDS)
constexpr int kTrsvMaxOrder = 8192;

// outer blocks of order N (row maps per system of LuWorkspace::pi)
inline int lu_solve_blocks(int N) { return (N + kLuOuter - 1) / kLuOuter; }
// floats of LuWorkspace::dinv per system
inline size_t lu_dinv_floats(int N) { return (size_t)N * 64 * 2 * 2; }

// Synchronisation state of the multi-workgroup triangular solves of one
// stream (orders <= kTrsvMaxOrder): ctx[4] = {ticket, finished blocks,
// epoch, error}; y and z [V][N][2] epoch-tagged words of the forward and
// backward values.  Set up once by launch_trsv_init (y and z cleared over
// n_words words each); a context serves one stream at a time (launches on it
// are ordered).
struct TrsvSync {
    uint32_t *ctx;
    uint64_t *y;
    uint64_t *z;
};
inline size_t trsv_words(int V, int N) { return (size_t)V * N * 2; }
void launch_trsv_init(const TrsvSync &sync, size_t n_words, hipStream_t stream);

// sizes of LuWorkspace::scratch_i (ints) and scratch_f (floats) per system
constexpr size_t kLuScratchI = kLuMaxOrder / 32;
constexpr size_t kLuScratchF = (32 * 32 + 64 * 32) * 2;

// 16-bit values of LuWorkspace::pack per system (four planes of N x kLuOuter)
inline size_t lu_pack_u16(int N) { return (size_t)N * kLuOuter * 4; }

// ints of LuWorkspace::gl per system
inline size_t lu_gather_ints(int N) { return (size_t)(N / 32) * kLuGather; }

// Pads rows/cols n..N-1 with the identity (zeros elsewhere).
void launch_pad_identity(float *A, int64_t lda, int64_t a_stride, int n, int N, int B,
                         hipStream_t stream);

// The same with per-system true orders n_sys[B] (device).
void launch_pad_identity_var(float *A, int64_t lda, int64_t a_stride, const int32_t *n_sys, int N,
                             int B, hipStream_t stream);

// 1-norm of each matrix (max column sum of |a_ij|) -> norms[B] and the
// largest modulus -> amax[B] (double).
void launch_norm1(const float *A, int64_t lda, int64_t a_stride, int n, int B, double *norms,
                  double *amax, hipStream_t stream);

// scale[b] = 2^-e with 2^e ~ amax[b] (operand scale of the FP16 update).
void launch_pow2_scale(const double *amax, float *scale, int B, hipStream_t stream);

// precision of the rank-kLuOuter trailing updates: 0 FP32, 1 FP16 (with
// scale[B]), 2 BF16; everything else runs in FP32.
void lu_factor(const LuWorkspace &w, int precision, const float *scale, hipStream_t stream);

// Solves A x = b (trans = 0) or A^H x = b (trans = 2) in place on x
// [B * vpm][ldx] interleaved complex (vector v against matrix v / vpm), using
// the LU factors and pivots.  With sync (capacity for B * vpm vectors) and
// the workspace's row maps, orders <= kTrsvMaxOrder run on N / 64
// workgroups per vector.
void lu_solve(const LuWorkspace &w, float *x, int64_t ldx, int trans, hipStream_t stream,
              int vpm = 1, const TrsvSync *sync = nullptr);

// Reciprocal condition estimate (cgecon / clacn2) of B factored systems,
// driven by the host: init X, then rounds of {copy X -> Y, lu_solve(Y,
// trans), step(trans)} for trans 0 and 2 until every system's state word
// state[4 b + 1] is -1; est[b] ends as the estimate of ||A_b^{-1}||_1.
void launch_rcond_init(float *X, int64_t ldx, int n, int N, int B, int32_t *state, double *est,
                       hipStream_t stream);
void launch_rcond_copy(const float *X, float *Y, int64_t ld, int N, int B, hipStream_t stream);
void launch_rcond_step(const float *Y, float *X, int64_t ldx, int n, int N, int B, int trans,
                       int32_t *state, double *est, hipStream_t stream);

// Residual r = b - A x in FP64 with A the FP32 operator (column-major
// interleaved), x and r FP64 interleaved.  Also writes per-system
// ||r||_inf and ||b||_inf into norms[2*b], norms[2*b+1].
// (B vectors, vector v against matrix v / vpm)
void launch_residual_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B,
                         const float *b, int64_t ldb, const double *x, int64_t ldx, double *r,
                         double *norms, hipStream_t stream, int vpm = 1);

// Helpers for refinement and rcond estimation.
void launch_copy_f32_to_f64(const float *src, double *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream);
void launch_copy_f64_to_f32(const double *src, float *dst, int64_t ld_src, int64_t ld_dst, int n,
                            int B, hipStream_t stream);
void launch_axpy_f32_to_f64(const float *d, double *x, int64_t ld_d, int64_t ld_x, int n, int B,
                            const int32_t *active, hipStream_t stream);

// Batched FP64 Krylov helpers (vectors interleaved complex double, stride
// ld per system; a basis V holds vector i of system b at (b * vstride +
// i * ldv) complex elements).
void launch_gemv_f64(const float *A, int64_t lda, int64_t a_stride, int n, int B, const double *x,
                     int64_t ldx, double *y, int64_t ldy, hipStream_t stream,
                     int vpm = 1);  // y = A x, B vectors, matrix v / vpm
void launch_dots_f64(const double *V, int64_t ldv, int64_t vstride, int m, const double *w,
                     int64_t ldw, int n, int B, double *h, int64_t ldh,
                     hipStream_t stream);  // h_i = V_i^H w, i < m
void launch_combine_f64(const double *V, int64_t ldv, int64_t vstride, int m, const double *c,
                        int64_t ldc, double sign, double *w, int64_t ldw, int n, int B,
                        const int32_t *active, hipStream_t stream);  // w += sign V c
void launch_norm2sq_f64(const double *w, int64_t ldw, int n, int B, double *out,
                        hipStream_t stream);
void launch_scale_f64(const double *src, int64_t lds, double *dst, int64_t ldd, int n, int B,
                      const double *f, hipStream_t stream);  // dst = f[b] src

// Copy a row-major split (re, im) host-layout matrix block on device into
// the column-major interleaved LU layout and back.
void launch_rowmajor_split_to_colmajor(const float *re, const float *im, int n, int B, float *A,
                                       int64_t lda, int64_t a_stride, hipStream_t stream);
void launch_colmajor_to_rowmajor_split(const float *A, int64_t lda, int64_t a_stride, int n,
                                       int B, float *re, float *im, hipStream_t stream);

}  // namespace hbem
