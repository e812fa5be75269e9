/*
 * hbem.h -- C ABI of libhbem, the MI355X (gfx950) per-frequency BEM hot path
 * behind hornlab_metal_bem's solve()/native_config() API.
 *
 * The reference crosses this boundary as a process + file IPC to a Swift/
 * Metal helper (hornlab_metal_bem/metal/native.py:233-1879, contract in
 * docs/native-ipc.md).  libhbem replaces that helper with an in-process
 * library: every entry point below names the reference call it stands in
 * for.  Conventions:
 *   - plain pointers and sizes only; arrays are C-contiguous little-endian;
 *   - the caller owns every host buffer it passes; the library never frees
 *     caller memory and never keeps a caller pointer past the call;
 *   - a session owns its device memory (geometry, singular-pair plan,
 *     quadrature tables, per-batch workspaces) until hbem_session_destroy;
 *   - every function returns 0 on success or a negative HBEM_E* code; the
 *     message of the last failure on the calling thread is returned by
 *     hbem_last_error();
 *   - calls on distinct sessions may run concurrently; calls on one session
 *     are serialised by the session.
 */
#ifndef HBEM_H
#define HBEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBEM_ABI_VERSION 6

#define HBEM_OK 0
#define HBEM_EINVAL -1      /* bad argument / contract violation */
#define HBEM_ENODEV -2      /* no usable gfx950 device */
#define HBEM_EHIP -3        /* HIP runtime error */
#define HBEM_ENOMEM -4      /* device allocation failed */
#define HBEM_ESINGULAR -5   /* LU hit an exactly zero pivot (lapack_info > 0) */
#define HBEM_EUNSUPPORTED -6

/* Symmetry codes: MS:697-717 (Geometry.symmetryPlaneCode). */
#define HBEM_SYM_NONE 0
#define HBEM_SYM_YZ 1
#define HBEM_SYM_XZ 2
#define HBEM_SYM_YZ_XZ 3
#define HBEM_SYM_XY 4

/* Geometry buffers in the layout of the reference's MetalGeometryBuffers
 * (hornlab_metal_bem/metal/geometry.py:36-65), already validated and
 * zero-based. */
typedef struct hbem_geometry {
    int32_t n_vertices;
    int32_t n_triangles;
    int32_t p1_dof_count;
    int32_t symmetry;               /* HBEM_SYM_* */
    const float *vertices_3xn;      /* [3][n_vertices] */
    const int32_t *triangles_3xm;   /* [3][n_triangles] */
    const int32_t *p1_local2global; /* [n_triangles][3] */
    const float *areas;             /* [n_triangles] */
    const float *normals_3xm;       /* [3][n_triangles] */
    const int32_t *physical_tags;   /* [n_triangles] */
} hbem_geometry;

/* Singular-pair plan of a session; the fields the reference reports under
 * duffy_corrections.planned_pairs (MS:610-640, MS:2195-2254). */
typedef struct hbem_session_info {
    int64_t duffy_coincident;
    int64_t duffy_edge;
    int64_t duffy_vertex;
    int64_t duffy_image_pairs;      /* pairs with a non-zero mask */
    int64_t duffy_unique_slots;     /* distinct (row, col) entries touched */
    int32_t n_triangles;
    int32_t p1_dof_count;
    int32_t symmetry;
    int32_t device_cu_count;
    double setup_seconds;           /* resident_context_seconds analogue */
    char device_name[64];
} hbem_session_info;

/* Dense-solve precision: dense_solve_dtype of config.py:236-246 plus the
 * mixed-precision modes of this library. */
#define HBEM_SOLVE_F32 0        /* FP32 complex LU (reference cgesv) */
#define HBEM_SOLVE_F32_IR 1     /* FP32 LU + FP64-residual refinement */
#define HBEM_SOLVE_MIXED_F16 2  /* FP16 MFMA trailing update + refinement */
#define HBEM_SOLVE_MIXED_BF16 3 /* BF16 MFMA trailing update + refinement */
/* dense_solve_dtype "float64" (the reference's zgesv of the FP32-assembled
 * system, main.swift:7101-7185): FP32 LU + GMRES-IR against the FP64
 * residual, run to ||b - A x||_inf / ||b||_inf <= min(refine_tolerance,
 * 1e-13) or until a restart brings no progress; a system left above 1e-10
 * reports refine_stalled. */
#define HBEM_SOLVE_F64 4

typedef struct hbem_solve_options {
    int32_t solve_mode;            /* HBEM_SOLVE_* */
    int32_t refine_max_iterations; /* refinement sweeps (0 = none) */
    double refine_tolerance;       /* stop at ||b-Ax||/||b|| <= tol */
    int32_t batch_frequencies;     /* systems resident at once (0 = auto) */
    int32_t estimate_rcond;        /* 1: cgecon-style rcond estimate */
    int32_t n_tags;                /* tags reduced into surface_pressure_avg */
    const int32_t *tags;           /* [n_tags] */
    int32_t impedance_tag;         /* tag of the impedance reduction */
    /* Robin walls (impedance_sources, config.py:174-196): normalised surface
     * admittance beta = rho c / Z_s per triangle, interleaved complex float
     * [F][n_triangles] (host memory for hbem_solve_batch, device memory for
     * hbem_solve_batch_device), or NULL for none.  The matrix gets
     * -(i k beta / 3) V on the triangle's three columns (main.swift:4101-4110)
     * and the field uses q + i k beta mean(p) (main.swift:1646-1674). */
    const float *robin_beta;
    /* right-hand sides per frequency sharing one matrix and one LU
     * (solve_multi_source, __init__.py:205-252); 0 or 1: one */
    int32_t n_sources;
    /* 1: capture the padding / norm / LU / first-solve launches of a resident
     * batch in a hipGraph and replay it for later batches of the same shape
     * and buffers (launch-bound small systems); 0: plain stream launches */
    int32_t lu_graph;
    /* CHIEF interior points (chief_points, config.py:212-217): [3][n_chief]
     * float (host memory for hbem_solve_batch, device memory for
     * hbem_solve_batch_device) or NULL.  Each frequency is then solved as the
     * least-squares problem min |A p - b|^2 + |s (C p - d)|^2 with
     * s = chief_weight |A|_inf / |C|_inf (solveDenseLeastSquaresZgels,
     * main.swift:7220-7416), with FP64-residual refinement. */
    const float *chief_points;
    int32_t n_chief;
    /* 1: regular-pair distances and phases in FP64 from vertex-anchored
     * offsets (FP64-grade matrices for shape gradients and accuracy
     * studies, about 1.5x the regular assembly time); 0: FP32.  The shape
     * adjoint's forward always uses 1.  No reference counterpart. */
    int32_t precise_geometry;
    double chief_weight;  /* > 0 (0 means 1) */
    /* hbem_dense_solve only: 0 solves A x = b, 2 solves A^H x = b (cgetrs
     * with trans 'C'; solve_mode HBEM_SOLVE_F32, no refinement); the batch
     * entry points need 0. */
    int32_t solve_trans;
    int32_t reserved[3];
} hbem_solve_options;

/* Per-frequency diagnostics: the case fields the reference returns
 * (native.py:1205-1320, keys native.py:2082-2137). */
typedef struct hbem_case_diag {
    double assembly_s;                   /* whole assembly stage, per case */
    double dense_solve_s;
    double field_s;                      /* field + reductions */
    int32_t lapack_info;
    int32_t refine_iterations;
    double dense_solve_rcond;            /* < 0 when not estimated */
    double dense_solve_condition_1norm;  /* ||A||_1 */
    double refine_residual_rel;          /* ||b-Ax||_inf/||b||_inf, FP64 */
    double regular_assembly_s;           /* regular-pair kernel alone (HIP events) */
    double singular_assembly_s;          /* singular-pair kernel alone (HIP events) */
    int32_t batch_index;                 /* resident chunk the case ran in */
    int32_t batch_size;                  /* cases in that chunk (times are chunk / size) */
    double near_assembly_s;              /* near-field kernel alone (0 when off) */
    int64_t near_pairs;                  /* near (test image, trial image) pairs corrected */
    double chief_residual_rel;           /* max over sources |s (C p - d)|_2 / |b|_2; -1 without CHIEF */
    int32_t refine_stalled;              /* 1: refinement stopped above its tolerance */
    int32_t reserved_i;
    double rcond_s;                      /* rcond estimate, per case (HIP events) */
} hbem_case_diag;

/* ABI version of the loaded library. */
int32_t hbem_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *hbem_last_error(void);

/* Device probe: number of visible gfx950 devices (0 if none).  Stands in for
 * discover_native_runtime (native.py:95-179). */
int32_t hbem_device_count(void);

/* Sustained rate (operations per second over the whole device) of the
 * hardware transcendental pipe on the mix the assembly kernels issue (sine,
 * cosine, reciprocal square root in equal parts): the ceiling the regular
 * assembly's transcendental side is measured against.  No reference
 * counterpart (a device probe, like hbem_device_count). */
int32_t hbem_probe_transcendental(double *ops_per_second);

typedef struct hbem_session hbem_session;

/* Upload geometry and build the resident plan (incidence, singular pairs,
 * quadrature tables).  Replaces MetalNativeStandardSession.create_session
 * (native.py:252-377) and ResidentMetalContext.init (MS:5370-5507). */
int32_t hbem_session_create(const hbem_geometry *geometry, hbem_session **out);

/* Release every device buffer of the session.  Replaces close()
 * (native.py:1709-1716). */
int32_t hbem_session_destroy(hbem_session *session);

/* Near-field quadrature of close non-touching pairs for every later
 * assembly of the session: level 1 or 2 subdivides both triangles into
 * 4^level sub-triangles with the 6-point rule on each and adds
 * (subdivided - 6x6) blocks; pairs are those whose mirrored triangles have a
 * bounding-box gap and a centroid distance below threshold x the longer
 * longest edge.  level 0 turns it off.  Replaces the
 * HORNLAB_METAL_BEM_NATIVE_NEAR_QUADRATURE setting read by the helper
 * (requestedNearQuadratureConfig, main.swift:5024-5055; pair list
 * buildNearPairList main.swift:2471-2558; corrections
 * applyNearFieldCorrectionsCPU main.swift:3072-3175).  pair_count (may be
 * NULL) receives the number of near pairs. */
int32_t hbem_session_set_near_field(hbem_session *session, int32_t level, double threshold,
                                    int64_t *pair_count);

/* Quadrature of every later assembly of the session; stands in for the
 * reference's HORNLAB_METAL_BEM_NATIVE_ASSEMBLY_MODE setting, read by
 * assembleRegular (main.swift:8128-8310):
 *   HBEM_ASSEMBLY_CORRECTED (default, "corrected"): the 6x6 rule on
 *     non-touching pairs, Sauter-Schwab quadrature on touching pairs;
 *   HBEM_ASSEMBLY_REGULAR ("optimized"): the 6x6 rule on every pair,
 *     touching pairs included, quadrature points that coincide excluded.
 * Near-field quadrature applies in both.  hbem_adjoint_forward needs
 * HBEM_ASSEMBLY_CORRECTED. */
#define HBEM_ASSEMBLY_CORRECTED 0
#define HBEM_ASSEMBLY_REGULAR 1
int32_t hbem_session_set_assembly_mode(hbem_session *session, int32_t mode);

int32_t hbem_session_get_info(hbem_session *session, hbem_session_info *info);

/* Assemble the corrected standard-Neumann system at one wavenumber:
 * A = (K - M/2 - (i k beta / 3) V) 2^planes (row-major n x n, split re/im
 * float32) and rhs = 2^planes V q.  Replaces assemble_standard_neumann
 * (native.py:388-480; helper op MS:8128-8310 in "corrected" mode).
 * q_re/q_im: DP0 Neumann data [n_triangles]; beta_re/beta_im: Robin
 * admittance [n_triangles] or NULL. */
int32_t hbem_assemble(hbem_session *session, float k_real, float k_imag, const float *q_re,
                      const float *q_im, const float *beta_re, const float *beta_im, float *A_re,
                      float *A_im, float *rhs_re, float *rhs_im);

/* Assemble, solve and evaluate the exterior field for n_freq frequencies.
 * Replaces assemble_solve_evaluate_standard_neumann_batch
 * (native.py:737-1203; helper MS:8869-9878).  With S = opts->n_sources
 * (>= 1) right-hand sides per frequency:
 *   k_real/k_imag [F]: assembly wavenumber (k_imag >= 0 for complex_k);
 *   field_k [F]: real wavenumber of the field kernel;
 *   q_re/q_im [F][S][n_triangles]: DP0 Neumann rows;
 *   obs_3xn [3][n_obs]: observation points;
 * outputs (host):
 *   field_re/im [F][S][n_obs];
 *   p_re/im [F][S][p1_dof_count] or NULL (write_surface_pressure);
 *   impedance [F][S][2] or NULL (area-mean pressure on opts->impedance_tag);
 *   pavg [F][S][n_tags][2] or NULL;
 *   diag [F] or NULL. */
int32_t hbem_solve_batch(hbem_session *session, int32_t n_freq, const float *k_real,
                         const float *k_imag, const float *field_k, const float *q_re,
                         const float *q_im, int32_t n_obs, const float *obs_3xn,
                         const hbem_solve_options *opts, float *field_re, float *field_im,
                         float *p_re, float *p_im, double *impedance, double *pavg,
                         hbem_case_diag *diag);

/* hbem_solve_batch with every array in device memory of the session's GPU
 * (for callers that keep data resident, e.g. a PyTorch wrapper).  No
 * counterpart in the reference, whose helper only reads files.  With
 * S = opts->n_sources (>= 1):
 *   k_real, field_k [F], k_imag [F] or NULL (real k);
 *   q [F][S][n_triangles] interleaved complex float;
 *   obs_3xn [3][n_obs];
 *   field [F][S][n_obs] interleaved complex float (overwritten);
 *   p [F][S][p1_dof_count] interleaved complex float or NULL;
 *   tag_avg [F][S][1 + n_tags][2] double: impedance tag, then opts->tags;
 *   diag [F] host array or NULL.
 * Returns after the results are complete on the device. */
int32_t hbem_solve_batch_device(hbem_session *session, int32_t n_freq, const float *k_real,
                                const float *k_imag, const float *field_k, const float *q,
                                int32_t n_obs, const float *obs_3xn,
                                const hbem_solve_options *opts, float *field, float *p,
                                double *tag_avg, hbem_case_diag *diag);

/* Forward pass of the shape adjoint: hbem_solve_batch_device (device
 * pointers, one source per frequency, near-field off) that keeps every
 * frequency's matrix, LU factors, surface solution and Neumann data resident
 * in the session until the next call on it.  No counterpart in the reference
 * (whose solver has no derivatives); it serves the torch.autograd wrapper of
 * waveguide shape optimisation. */
int32_t hbem_adjoint_forward(hbem_session *session, int32_t n_freq, const float *k_real,
                             const float *k_imag, const float *field_k, const float *q,
                             int32_t n_obs, const float *obs_3xn,
                             const hbem_solve_options *options, float *field, float *p,
                             double *tag_avg, hbem_case_diag *diag);

/* Backward pass after hbem_adjoint_forward.  grad_field [F][n_obs]
 * interleaved complex float (device) holds dL/dRe f + i dL/dIm f of a real
 * loss L of the field.  Writes grad_vertices [3][n_vertices] double (device,
 * dL/d vertex xyz, overwritten) and, when non-NULL, grad_q [F][T] interleaved
 * complex float (device, dL/dRe q + i dL/dIm q).  The transpose solve reuses
 * the forward LU factors and takes refine_iterations FP64-residual
 * refinement steps. */
int32_t hbem_adjoint_backward(hbem_session *session, const float *grad_field,
                              double *grad_vertices, float *grad_q, int32_t refine_iterations);

/* Exterior field from given surface data.  Replaces
 * evaluate_standard_exterior_batch (native.py:1552-1707; kernel
 * MS:4135-4203).  p [F][p1_dof_count], q [F][n_triangles], out [F][n_obs]. */
int32_t hbem_evaluate_field(hbem_session *session, int32_t n_freq, const float *field_k,
                            const float *p_re, const float *p_im, const float *q_re,
                            const float *q_im, int32_t n_obs, const float *obs_3xn,
                            float *out_re, float *out_im);

/* Dense complex solve of caller-supplied systems on the GPU LU (row-major
 * split float32 A [F][n][n], b [F][n] -> x [F][n]).  Replaces the
 * Accelerate cgesv call site solveDenseAccelerateCgesvMulti (MS:7834-7941).
 * diag may be NULL. */
int32_t hbem_dense_solve(int32_t n_sys, int32_t n, const float *A_re, const float *A_im,
                         const float *b_re, const float *b_im, const hbem_solve_options *opts,
                         float *x_re, float *x_im, hbem_case_diag *diag);

/* Session groups: the sweeps of several meshes solved together (the
 * batched uneven sweep; the reference runs such cases through its resident
 * context's double-buffered case slots and solve pool, main.swift:5520-5562,
 * 9332-9832).  One regular and one singular assembly launch cover every
 * member's tile pairs and singular classes; the members' systems are padded
 * (identity) into size classes of a common order, each class factored and
 * solved as one batch on its own stream; field and reductions per member;
 * the whole sequence is captured once per shape and buffer set and replayed
 * as one hipGraph.  Members must share the device and the symmetry image
 * count, use the corrected assembly and have near-field quadrature off; the
 * group borrows the sessions (destroy the group first). */
typedef struct hbem_group hbem_group;

/* One member's arrays for hbem_group_solve_device: device memory, the
 * meanings of hbem_solve_batch_device's arguments with one source. */
typedef struct hbem_group_case {
    int32_t n_freq;
    const float *k_real;   /* [F] */
    const float *field_k;  /* [F] */
    const float *q;        /* [F][n_triangles] interleaved complex */
    int32_t n_obs;
    const float *obs_3xn;  /* [3][n_obs] */
    float *field;          /* [F][n_obs] interleaved complex (overwritten) */
    float *p;              /* [F][p1_dof_count] interleaved complex or NULL */
    double *tag_avg;       /* [F][1 + n_tags][2] */
} hbem_group_case;

int32_t hbem_group_create(int32_t n_sessions, hbem_session *const *sessions, hbem_group **out);
int32_t hbem_group_destroy(hbem_group *group);

/* Solves every member's frequencies (cases[n_sessions]).  opts: solve_mode
 * HBEM_SOLVE_F32 without rcond, Robin walls, CHIEF rows, extra sources or
 * precise geometry (HBEM_EUNSUPPORTED otherwise); tags as in the batch calls.
 * diag [sum of n_freq] (members in order); member_seconds [n_sessions] or
 * NULL: the time from the sweep's start until the member's results are
 * complete (its systems' latency); sweep_seconds or NULL: the sweep. */
int32_t hbem_group_solve_device(hbem_group *group, const hbem_group_case *cases,
                                const hbem_solve_options *opts, hbem_case_diag *diag,
                                double *member_seconds, double *sweep_seconds);

/* Times one sweep for each of eight launch shapes -- size-class step
 * (padded-order granularity 64, 256, 512, 1024: from one class per mesh to
 * a few large classes with more padding) x assembly as one launch over all
 * members or one launch per size class (whose solves then overlap the next
 * class's assembly) -- and keeps the fastest for later
 * hbem_group_solve_device calls.  granularity (or NULL) receives the chosen
 * step, negative when the assembly is split per class; seconds[8] (or NULL)
 * the timings, in the order (64, 256, 512, 1024) one launch, then split. */
int32_t hbem_group_autotune(hbem_group *group, const hbem_group_case *cases,
                            const hbem_solve_options *opts, int32_t *granularity,
                            double *seconds);

/* hbem_dense_solve on device arrays without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its working storage is a fixed device array of the library,
 * HBEM_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU
 * race (each call overwrites what the previous one left there).
 *   A_re, A_im [n_sys][n][n] row-major float, b_re, b_im [n_sys][n],
 *   x_re, x_im [n_sys][n] (written), info [n_sys] int32 or NULL (LAPACK
 *   info: 0, or the first column without a nonzero pivot): device memory of
 *   the current device;
 *   opts: NULL, or solve_mode HBEM_SOLVE_F32 with no rcond, refinement,
 *   CHIEF rows or extra sources (solve_trans 0 or 2 is honoured);
 *   stream: the hipStream_t to enqueue on (NULL: the null stream).
 * Limits: n_sys <= 64, n <= 2048, n_sys * N^2 <= 2^22 with N = n rounded up
 * to 64; larger problems are refused (HBEM_EINVAL) and take hbem_dense_solve.
 * The call only enqueues work (no synchronisation).  The first call on a
 * device looks up the array's address there (hipGetSymbolAddress): make it
 * outside graph capture. */
#define HBEM_NOALLOC_SCRATCH_BYTES 37257472
int32_t hbem_dense_solve_noalloc(int32_t n_sys, int32_t n, const float *A_re, const float *A_im,
                                 const float *b_re, const float *b_im,
                                 const hbem_solve_options *opts, float *x_re, float *x_im,
                                 int32_t *info, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HBEM_H */
