// C ABI of libhbem (include/hbem.h): sessions, device workspaces and the
// per-batch orchestration of assembly -> LU -> refinement -> field ->
// reductions on one HIP stream per session.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hbem.h"
#include "hbem_internal.h"

namespace {

thread_local std::string g_last_error;

struct HbemError : std::runtime_error {
    int code;
    HbemError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string &msg) { throw HbemError(code, msg); }

template <typename F>
int32_t guarded(F &&f) {
    try {
        g_last_error.clear();
        f();
        return HBEM_OK;
    } catch (const HbemError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return HBEM_ENOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return HBEM_EINVAL;
    }
}

void hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? HBEM_ENOMEM : HBEM_EHIP;
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

// Growable device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t want) {
        if (want <= bytes) return;
        if (ptr) hip_check(hipFree(ptr), "hipFree");
        ptr = nullptr;
        bytes = 0;
        hip_check(hipMalloc(&ptr, want), "hipMalloc");
        bytes = want;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(ptr);
    }
};

template <typename T>
void upload(DevBuf &d, const std::vector<T> &h, hipStream_t s) {
    d.reserve(std::max<size_t>(1, h.size() * sizeof(T)));
    if (!h.empty()) {
        hip_check(hipMemcpyAsync(d.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync H2D");
        // the host vector may be a temporary: finish the copy before returning
        hip_check(hipStreamSynchronize(s), "upload sync");
    }
}

int gfx950_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return -1;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) return d;
    }
    return -1;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer() {
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { hip_check(hipEventRecord(a, s), "hipEventRecord"); }
    void stop(hipStream_t s) { hip_check(hipEventRecord(b, s), "hipEventRecord"); }
    double seconds() {
        hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms * 1e-3;
    }
};

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

namespace hbem {
int32_t run_guarded(const std::function<void()> &f) { return guarded(f); }
void fail_with(int code, const std::string &msg) { fail(code, msg); }
void check_hip(hipError_t e, const char *what) { hip_check(e, what); }
}  // namespace hbem

struct hbem_session {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    hbem::Plan plan;
    hbem::DeviceGeometry dg{};
    DevBuf tri_f_aos, tri_i_aos, tri_t_f, tri_t_i, tri_f_soa, tri_i_soa, sg_f, sg_i, sg_rule, sg_rule_reg, tile_start,
        tile_count, tile_ndofs, tile_dofs, tile_pairs, nc_f, nc_i, nc_rule, sg_adj, vtx_d, triv_d,
        normals_d, jac_d;
    // CHIEF rows and solves of the current resident batch
    DevBuf chief_C, chief_d, chief_pts, chief_z, chief_e, chief_z64, chief_r, chief_ainf;
    const float *chief_pts_d = nullptr;
    std::vector<float> chief_b;       // original right-hand sides of the batch
    std::vector<double> chief_scale;  // s per matrix
    // adjoint store (hbem_adjoint_forward -> hbem_adjoint_backward)
    DevBuf adj_k, adj_kim, adj_fk, adj_q, adj_beta, adj_obs, adj_c, adj_ft, adj_sig, adj_gN,
        adj_l32, adj_e64, adj_l64, adj_r64, adj_act;
    // captured factor + solve sequence of the last resident batch shape
    hipGraphExec_t lu_graph = nullptr;
    std::vector<uintptr_t> lu_graph_key;
    bool adj_keep = false;   // forward in progress: one resident batch, A kept
    bool adj_valid = false;  // the store matches the session's workspaces
    int adj_F = 0, adj_nobs = 0;
    bool adj_complex = false, adj_robin = false;
    // per-call workspaces
    DevBuf A, A_orig, rhs, x, x64, r64, xprev, ipiv, info, gl, lu_scr_i, lu_scr_f, lu_pack, amax,
        lu_scale, kv, kh, kn, kf,
        q, kre, kim, fk, obs, field, tags,
        tagout, norms, anorm, active, rowsplit, rcx, rcy, rc_state, rc_est;
    // staging of host-input batches
    DevBuf io_k, io_kim, io_fk, io_q, io_obs, io_field, io_p, io_tag, io_beta, qt, l2g;
    // second stream of the session: the rcond iteration, concurrent with
    // the refinement and the field on `stream`
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_lu = nullptr;  // end of the factorization, for the aux stream
    // row maps of the factors and the solve contexts of the two streams
    // ([0] stream, [1] aux_stream) for the multi-workgroup triangular solves
    // (orders <= 8192); trsv_on while the current solve call uses them
    DevBuf lu_perm, lu_pi, lu_lf, lu_dinv, ts_ctx[2], ts_y[2], ts_z[2];
    hbem::TrsvSync trsv[2] = {};
    bool trsv_on = false;
    double setup_seconds = 0.0;
    int64_t near_pairs = 0;
    int assembly_mode = HBEM_ASSEMBLY_CORRECTED;
    std::string device_name;
    int cu_count = 0;

    ~hbem_session() {
        for (DevBuf *b : {&tri_f_aos, &tri_i_aos, &tri_t_f, &tri_t_i, &tri_f_soa, &tri_i_soa, &sg_f, &sg_i,
                          &sg_rule, &sg_rule_reg, &tile_start, &tile_count, &tile_ndofs, &tile_dofs, &tile_pairs, &nc_f, &nc_i, &nc_rule, &sg_adj, &vtx_d, &triv_d, &normals_d, &jac_d, &adj_k, &adj_kim, &adj_fk, &adj_q, &adj_beta, &adj_obs, &adj_c, &adj_ft, &adj_sig, &adj_gN, &adj_l32, &adj_e64, &adj_l64, &adj_r64, &adj_act, &chief_C, &chief_d, &chief_pts, &chief_z, &chief_e, &chief_z64, &chief_r, &chief_ainf, &A, &A_orig, &rhs, &x, &x64, &r64, &xprev, &ipiv, &info, &gl, &lu_scr_i, &lu_scr_f, &lu_pack, &amax, &lu_scale, &kv, &kh, &kn, &kf, &q, &kre,
                          &kim, &fk, &obs, &field, &tags, &tagout, &norms, &anorm, &active,
                          &rowsplit, &rcx, &rcy, &rc_state, &rc_est, &io_k, &io_kim, &io_fk, &io_q, &io_obs, &io_field,
                          &io_p, &io_tag, &io_beta, &qt, &l2g, &lu_perm, &lu_pi, &lu_lf, &lu_dinv, &ts_ctx[0],
                          &ts_ctx[1], &ts_y[0], &ts_y[1], &ts_z[0], &ts_z[1]})
            b->release();
        if (lu_graph) (void)hipGraphExecDestroy(lu_graph);
        if (aux_stream) (void)hipStreamDestroy(aux_stream);
        if (ev_lu) (void)hipEventDestroy(ev_lu);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void set_device(const hbem_session *s) { hip_check(hipSetDevice(s->device), "hipSetDevice"); }

// Interleave split (re, im) host arrays of length n (count rows of stride n).
std::vector<float> interleave(const float *re, const float *im, size_t n) {
    std::vector<float> out(2 * n);
    for (size_t i = 0; i < n; ++i) {
        out[2 * i] = re ? re[i] : 0.0f;
        out[2 * i + 1] = im ? im[i] : 0.0f;
    }
    return out;
}

// Assembles B systems (frequencies) into s->A (column-major, lda, stride)
// and s->rhs from device inputs: k_real/k_imag [B] (k_imag may be null for
// real k) and interleaved complex Neumann rows q [B][T].  Optional timers
// bracket the regular and the singular kernel on the session stream.
bool any_nonzero(const float *v, size_t n) {
    if (!v) return false;
    for (size_t i = 0; i < n; ++i)
        if (v[i] != 0.0f) return true;
    return false;
}

// B matrices (one per frequency) and B * S right-hand sides (q [B][S][T],
// rhs vector b * S + src), beta [B][T] Robin admittance or null.
void assemble_device(hbem_session *s, int B, int S, const float *kre_d, const float *kim_d,
                     const float *q_d, const float *beta_d, int64_t lda, int64_t a_stride,
                     int64_t rhs_stride, Timer *t_reg, Timer *t_sing, Timer *t_near = nullptr,
                     bool precise = false) {
    const size_t a_floats = (size_t)B * a_stride * 2;
    s->A.reserve(a_floats * sizeof(float));
    s->rhs.reserve((size_t)B * S * rhs_stride * 2 * sizeof(float));
    hbem::launch_fill_zero(s->A.as<float>(), a_floats, s->stream);
    hbem::launch_fill_zero(s->rhs.as<float>(), (size_t)B * S * rhs_stride * 2, s->stream);
    if (t_reg) t_reg->start(s->stream);
    hbem::launch_assemble_regular(s->dg, B, kre_d, kim_d, q_d, S, beta_d, s->A.as<float>(), lda,
                                  a_stride, s->rhs.as<float>(), rhs_stride, s->stream, precise);
    if (t_reg) t_reg->stop(s->stream);
    if (t_sing) t_sing->start(s->stream);
    hbem::launch_assemble_singular(s->dg, B, kre_d, kim_d, q_d, S, beta_d, s->A.as<float>(), lda,
                                   a_stride, s->rhs.as<float>(), rhs_stride, s->stream);
    if (t_sing) t_sing->stop(s->stream);
    if (t_near) t_near->start(s->stream);
    hbem::launch_assemble_near(s->dg, B, kre_d, kim_d, q_d, S, beta_d, s->A.as<float>(), lda,
                               a_stride, s->rhs.as<float>(), rhs_stride, s->stream);
    if (t_near) t_near->stop(s->stream);
    hip_check(hipGetLastError(), "assembly launch");
}

// Host-input twin of assemble_device (uploads k and split q rows).
void assemble_batch(hbem_session *s, int B, const float *k_real_h, const float *k_imag_h,
                    const float *q_re_h, const float *q_im_h, const float *beta_re_h,
                    const float *beta_im_h, int64_t lda, int64_t a_stride, int64_t rhs_stride,
                    bool complex_k) {
    const int T = s->plan.n_triangles;
    std::vector<float> kr(k_real_h, k_real_h + B);
    upload(s->kre, kr, s->stream);
    if (complex_k) {
        std::vector<float> ki(k_imag_h, k_imag_h + B);
        upload(s->kim, ki, s->stream);
    }
    std::vector<float> q = interleave(q_re_h, q_im_h, (size_t)B * T);
    upload(s->q, q, s->stream);
    const bool robin = any_nonzero(beta_re_h, (size_t)B * T) || any_nonzero(beta_im_h, (size_t)B * T);
    if (robin) {
        std::vector<float> bz((size_t)B * T, 0.0f);
        upload(s->io_beta, interleave(beta_re_h ? beta_re_h : bz.data(),
                                      beta_im_h ? beta_im_h : bz.data(), (size_t)B * T),
               s->stream);
    }
    assemble_device(s, B, 1, s->kre.as<float>(), complex_k ? s->kim.as<float>() : nullptr,
                    s->q.as<float>(), robin ? s->io_beta.as<float>() : nullptr, lda, a_stride,
                    rhs_stride, nullptr, nullptr);
}


// Reciprocal condition number of each system of a factored batch, as the
// reference's cgecon call (estimateReciprocalCondition, main.swift:475-503):
// ||A^{-1}||_1 by the clacn2 iteration (Hager / Higham), one state machine per
// system on the device, every round one batched A^{-1} and/or A^{-H} solve of
// the systems that need it; the host only reads the B solve flags per round.
// rcond = 1 / (||A||_1 ||A^{-1}||_1), 0 for a zero estimate or norm.
void reserve_rcond(hbem_session *s, int B, int N) {
    s->rcx.reserve((size_t)B * N * 2 * sizeof(float));
    s->rcy.reserve((size_t)B * N * 2 * sizeof(float));
    s->rc_state.reserve(sizeof(int32_t) * 4 * (size_t)B);
    s->rc_est.reserve(sizeof(double) * (size_t)B);
}

// Solve context `which` (0 for the session stream, 1 for the aux stream) with
// room for V vectors of order N; initialised on `st` whenever its buffers
// move (outside any graph capture: the context's epoch then advances on the
// device, launch by launch).
hbem::TrsvSync trsv_sync(hbem_session *s, int which, int V, int N, hipStream_t st) {
    DevBuf &c = s->ts_ctx[which], &y = s->ts_y[which], &z = s->ts_z[which];
    const void *c0 = c.ptr, *y0 = y.ptr, *z0 = z.ptr;
    const size_t words = hbem::trsv_words(V, N);
    c.reserve(sizeof(uint32_t) * 4);
    y.reserve(sizeof(uint64_t) * words);
    z.reserve(sizeof(uint64_t) * words);
    const hbem::TrsvSync sync{c.as<uint32_t>(), y.as<uint64_t>(), z.as<uint64_t>()};
    if (c.ptr != c0 || y.ptr != y0 || z.ptr != z0)
        hbem::launch_trsv_init(sync, std::min(y.bytes, z.bytes) / sizeof(uint64_t), st);
    return sync;
}

// Runs on `st`; its buffers come from reserve_rcond (no allocation here, so
// it can run beside other work of the session on another host thread).
// Returns the estimates of ||A^{-1}||_1; rcond = 1 / (||A||_1 est).
void estimate_rcond(hbem_session *s, const hbem::LuWorkspace &w, std::vector<double> *est_out,
                    hipStream_t st) {
    const int B = w.B, n = w.n, N = w.N;
    float *X = s->rcx.as<float>(), *Y = s->rcy.as<float>();
    int32_t *state = s->rc_state.as<int32_t>();
    double *est_d = s->rc_est.as<double>();
    hbem::launch_rcond_init(X, N, n, N, B, state, est_d, st);
    std::vector<int32_t> sw(4 * (size_t)B);
    // every system needs an A^{-1} solve first; the iteration takes at most
    // 2 + 2 * 4 + 1 solves per system (clacn2's itmax = 5)
    bool need[3] = {true, false, false};
    for (int round = 0; round < 16 && (need[0] || need[2]); ++round) {
        for (int trans : {0, 2}) {
            if (!need[trans]) continue;
            hbem::launch_rcond_copy(X, Y, N, N, B, st);
            hbem::lu_solve(w, Y, N, trans, st, 1, s->trsv_on ? &s->trsv[1] : nullptr);
            hbem::launch_rcond_step(Y, X, N, n, N, B, trans, state, est_d, st);
        }
        hip_check(hipGetLastError(), "rcond launch");
        hip_check(hipMemcpyAsync(sw.data(), state, sizeof(int32_t) * sw.size(),
                                 hipMemcpyDeviceToHost, st),
                  "rcond state D2H");
        hip_check(hipStreamSynchronize(st), "rcond sync");
        need[0] = need[2] = false;
        for (int b = 0; b < B; ++b) {
            const int op = sw[4 * (size_t)b + 1];
            if (op == 0 || op == 2) need[op] = true;
        }
    }
    est_out->assign(B, 0.0);
    hip_check(hipMemcpyAsync(est_out->data(), est_d, sizeof(double) * B, hipMemcpyDeviceToHost,
                             st),
              "rcond D2H");
    hip_check(hipStreamSynchronize(st), "rcond sync");
}

}  // namespace

extern "C" {

int32_t hbem_abi_version(void) { return HBEM_ABI_VERSION; }

const char *hbem_last_error(void) { return g_last_error.c_str(); }

int32_t hbem_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    int n = 0;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++n;
    }
    return n;
}

int32_t hbem_probe_transcendental(double *ops_per_second) {
    return guarded([&] {
        if (!ops_per_second) fail(HBEM_EINVAL, "hbem_probe_transcendental: null argument");
        const int dev = gfx950_device();
        if (dev < 0) fail(HBEM_ENODEV, "no gfx950 (MI355X) device is visible");
        hip_check(hipSetDevice(dev), "hipSetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        // 8 blocks of 256 per CU (8 waves per SIMD), long enough to time
        const int blocks = prop.multiProcessorCount * 8, iters = 4096;
        hipStream_t st = nullptr;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        DevBuf out;
        double best = 0.0;
        try {
            out.reserve(sizeof(float) * (size_t)blocks * 256);
            hbem::launch_probe_transcendental(out.as<float>(), blocks, 64, st);  // warm-up
            for (int rep = 0; rep < 3; ++rep) {
                Timer t;
                t.start(st);
                hbem::launch_probe_transcendental(out.as<float>(), blocks, iters, st);
                t.stop(st);
                hip_check(hipGetLastError(), "probe launch");
                const double ops = (double)blocks * 256.0 * iters * 16.0;
                best = std::max(best, ops / t.seconds());
            }
        } catch (...) {
            out.release();
            (void)hipStreamDestroy(st);
            throw;
        }
        out.release();
        (void)hipStreamDestroy(st);
        *ops_per_second = best;
    });
}

int32_t hbem_session_create(const hbem_geometry *geom, hbem_session **out) {
    return guarded([&] {
        if (!geom || !out) fail(HBEM_EINVAL, "hbem_session_create: null argument");
        *out = nullptr;
        if (!geom->vertices_3xn || !geom->triangles_3xm || !geom->p1_local2global ||
            !geom->areas || !geom->normals_3xm)
            fail(HBEM_EINVAL, "hbem_session_create: geometry arrays must be non-null");
        const double t0 = now_s();
        const int dev = gfx950_device();
        if (dev < 0) fail(HBEM_ENODEV, "no gfx950 (MI355X) device is visible");
        auto s = std::make_unique<hbem_session>();
        s->device = dev;
        set_device(s.get());
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        s->device_name = prop.name;
        s->cu_count = prop.multiProcessorCount;
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        const std::string err = hbem::build_plan(
            geom->n_vertices, geom->n_triangles, geom->p1_dof_count, geom->symmetry,
            geom->vertices_3xn, geom->triangles_3xm, geom->p1_local2global, geom->areas,
            geom->normals_3xm, geom->physical_tags, &s->plan);
        if (!err.empty()) fail(HBEM_EINVAL, "hbem_session_create: " + err);
        const hbem::Plan &P = s->plan;
        upload(s->tri_f_aos, P.tri_f_aos, s->stream);
        upload(s->tri_i_aos, P.tri_i_aos, s->stream);
        upload(s->tri_t_f, P.tri_t_f, s->stream);
        upload(s->tri_t_i, P.tri_t_i, s->stream);
        upload(s->tri_f_soa, P.tri_f_soa, s->stream);
        upload(s->tri_i_soa, P.tri_i_soa, s->stream);
        upload(s->sg_f, P.sg_f, s->stream);
        upload(s->sg_i, P.sg_i, s->stream);
        upload(s->sg_rule, P.sg_rule, s->stream);
        upload(s->sg_rule_reg, P.sg_rule_reg, s->stream);
        upload(s->tile_start, P.tile_start, s->stream);
        upload(s->tile_count, P.tile_count, s->stream);
        upload(s->tile_ndofs, P.tile_ndofs, s->stream);
        upload(s->tile_dofs, P.tile_dofs, s->stream);
        upload(s->tile_pairs, P.tile_pairs, s->stream);
        upload(s->l2g, P.l2g, s->stream);
        upload(s->sg_adj, P.sg_adj, s->stream);
        upload(s->vtx_d, P.vtx, s->stream);
        upload(s->triv_d, P.tri, s->stream);
        upload(s->normals_d, P.normals, s->stream);
        {
            std::vector<float> jac(P.areas.size());
            for (size_t t = 0; t < jac.size(); ++t) jac[t] = 2.0f * P.areas[t];
            upload(s->jac_d, jac, s->stream);
            hip_check(hipStreamSynchronize(s->stream), "session upload");
        }
        hbem::DeviceGeometry &g = s->dg;
        g.tri_f_aos = s->tri_f_aos.as<float>();
        g.tri_i_aos = s->tri_i_aos.as<int32_t>();
        g.tri_t_f = s->tri_t_f.as<float>();
        g.tri_t_i = s->tri_t_i.as<int32_t>();
        g.tri_f_soa = s->tri_f_soa.as<float>();
        g.tri_i_soa = s->tri_i_soa.as<int32_t>();
        g.sg_f = s->sg_f.as<float>();
        g.sg_i = s->sg_i.as<int32_t>();
        g.sg_rule = s->sg_rule.as<float>();
        g.sg_np = P.sg_np;
        for (int i = 0; i < 3; ++i) {
            g.sg_waves[i] = P.sg_waves[i];
            g.sg_rule_offset[i] = P.sg_rule_offset[i];
            g.sg_rule_count[i] = P.sg_rule_count[i];
        }
        g.n_triangles = P.n_triangles;
        g.n_dofs = P.n_dofs;
        g.symmetry = P.symmetry;
        g.n_masks = P.n_masks;
        for (int i = 0; i < hbem::kMaxMasks; ++i) g.masks[i] = P.masks[i];
        g.row_weight = P.row_weight;
        g.tile_start = s->tile_start.as<int32_t>();
        g.tile_count = s->tile_count.as<int32_t>();
        g.tile_ndofs = s->tile_ndofs.as<int32_t>();
        g.tile_dofs = s->tile_dofs.as<int32_t>();
        g.tile_pairs = s->tile_pairs.as<int32_t>();
        g.l2g = s->l2g.as<int32_t>();
        g.sg_adj = s->sg_adj.as<int32_t>();
        g.vtx = s->vtx_d.as<float>();
        g.tri_v = s->triv_d.as<int32_t>();
        g.normals = s->normals_d.as<float>();
        g.jac_c = s->jac_d.as<float>();
        g.n_vertices = P.n_vertices;
        g.n_tile_pairs = (int)(P.tile_pairs.size() / 2);
        g.n_tiles = (int)P.tile_start.size();
        g.max_tile_dofs = P.max_tile_dofs;
        hip_check(hipStreamSynchronize(s->stream), "session upload");
        s->setup_seconds = now_s() - t0;
        *out = s.release();
    });
}

int32_t hbem_session_set_near_field(hbem_session *s, int32_t level, double threshold,
                                    int64_t *pair_count) {
    return guarded([&] {
        if (!s) fail(HBEM_EINVAL, "hbem_session_set_near_field: null session");
        if (level < 0 || level > 2)
            fail(HBEM_EINVAL, "hbem_session_set_near_field: level must be 0, 1 or 2");
        if (level > 0 && !(std::isfinite(threshold) && threshold > 0.0))
            fail(HBEM_EINVAL, "hbem_session_set_near_field: threshold must be finite and positive");
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        set_device(s);
        hbem::DeviceGeometry &g = s->dg;
        if (level == 0) {
            g.nc_np = 0;
            g.nc_sub_points = 0;
            g.nc_f = nullptr;
            g.nc_i = nullptr;
            g.nc_rule = nullptr;
            s->near_pairs = 0;
        } else if (s->plan.nc_level != level || s->plan.nc_threshold != threshold || g.nc_rule == nullptr) {
            hbem::Plan &P = s->plan;
            hbem::build_near_classes(&P, threshold, level);
            hip_check(hipStreamSynchronize(s->stream), "near-field upload");
            upload(s->nc_f, P.nc_f, s->stream);
            upload(s->nc_i, P.nc_i, s->stream);
            upload(s->nc_rule, P.nc_rule, s->stream);
            hip_check(hipStreamSynchronize(s->stream), "near-field upload");
            g.nc_f = s->nc_f.as<float>();
            g.nc_i = s->nc_i.as<int32_t>();
            g.nc_rule = s->nc_rule.as<float>();
            g.nc_np = P.nc_np;
            g.nc_sub_points = P.nc_sub_points;
            s->near_pairs = P.nc_pairs;
        }
        if (pair_count) *pair_count = s->near_pairs;
    });
}

int32_t hbem_session_set_assembly_mode(hbem_session *s, int32_t mode) {
    return guarded([&] {
        if (!s) fail(HBEM_EINVAL, "hbem_session_set_assembly_mode: null session");
        if (mode != HBEM_ASSEMBLY_CORRECTED && mode != HBEM_ASSEMBLY_REGULAR)
            fail(HBEM_EINVAL, "hbem_session_set_assembly_mode: unknown mode");
        std::lock_guard<std::mutex> lock(s->mu);
        const hbem::Plan &P = s->plan;
        const bool reg = mode == HBEM_ASSEMBLY_REGULAR;
        s->dg.sg_rule = (reg ? s->sg_rule_reg : s->sg_rule).as<float>();
        for (int i = 0; i < 3; ++i) {
            s->dg.sg_rule_offset[i] = reg ? P.sg_rule_reg_offset[i] : P.sg_rule_offset[i];
            s->dg.sg_rule_count[i] = reg ? P.sg_rule_reg_count[i] : P.sg_rule_count[i];
        }
        s->assembly_mode = mode;
        s->adj_valid = false;
    });
}

int32_t hbem_session_destroy(hbem_session *session) {
    return guarded([&] {
        if (!session) return;
        (void)hipSetDevice(session->device);
        (void)hipStreamSynchronize(session->stream);
        delete session;
    });
}

int32_t hbem_session_get_info(hbem_session *s, hbem_session_info *info) {
    return guarded([&] {
        if (!s || !info) fail(HBEM_EINVAL, "hbem_session_get_info: null argument");
        std::memset(info, 0, sizeof(*info));
        info->duffy_coincident = s->plan.kind_count[0];
        info->duffy_edge = s->plan.kind_count[1];
        info->duffy_vertex = s->plan.kind_count[2];
        info->duffy_image_pairs = s->plan.image_pairs;
        info->duffy_unique_slots = s->plan.unique_slots;
        info->n_triangles = s->plan.n_triangles;
        info->p1_dof_count = s->plan.n_dofs;
        info->symmetry = s->plan.symmetry;
        info->device_cu_count = s->cu_count;
        info->setup_seconds = s->setup_seconds;
        std::snprintf(info->device_name, sizeof(info->device_name), "%s", s->device_name.c_str());
    });
}

int32_t hbem_assemble(hbem_session *s, float k_real, float k_imag, const float *q_re,
                      const float *q_im, const float *beta_re, const float *beta_im, float *A_re,
                      float *A_im, float *rhs_re, float *rhs_im) {
    return guarded([&] {
        if (!s) fail(HBEM_EINVAL, "hbem_assemble: null session");
        if (!A_re || !A_im || !rhs_re || !rhs_im)
            fail(HBEM_EINVAL, "hbem_assemble: output arrays must be non-null");
        if (!std::isfinite(k_real) || !std::isfinite(k_imag) || k_imag < 0.0f)
            fail(HBEM_EINVAL, "hbem_assemble: k must be finite with k_imag >= 0");
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        set_device(s);
        const int n = s->plan.n_dofs;
        const int64_t lda = n, a_stride = (int64_t)n * n, rhs_stride = n;
        assemble_batch(s, 1, &k_real, &k_imag, q_re, q_im, beta_re, beta_im, lda, a_stride,
                       rhs_stride, k_imag != 0.0f);
        s->rowsplit.reserve((size_t)2 * n * n * sizeof(float));
        float *re_d = s->rowsplit.as<float>();
        float *im_d = re_d + (size_t)n * n;
        hbem::launch_colmajor_to_rowmajor_split(s->A.as<float>(), lda, a_stride, n, 1, re_d, im_d,
                                                s->stream);
        hip_check(hipMemcpyAsync(A_re, re_d, sizeof(float) * (size_t)n * n, hipMemcpyDeviceToHost,
                                 s->stream),
                  "A_re D2H");
        hip_check(hipMemcpyAsync(A_im, im_d, sizeof(float) * (size_t)n * n, hipMemcpyDeviceToHost,
                                 s->stream),
                  "A_im D2H");
        std::vector<float> rh((size_t)2 * n);
        hip_check(hipMemcpyAsync(rh.data(), s->rhs.ptr, sizeof(float) * 2 * (size_t)n,
                                 hipMemcpyDeviceToHost, s->stream),
                  "rhs D2H");
        hip_check(hipStreamSynchronize(s->stream), "hbem_assemble");
        for (int i = 0; i < n; ++i) {
            rhs_re[i] = rh[2 * i];
            rhs_im[i] = rh[2 * i + 1];
        }
    });
}

}  // extern "C"

namespace {

// Target relative residual ||b - A x||_inf / ||b||_inf of the refinement:
// the caller's tolerance; for the FP64-accurate mode at most 1e-13, so the
// iteration runs to the FP64 residual floor (it stops earlier when a restart
// brings no progress, as LAPACK's refinement does).
double refine_tolerance(const hbem_solve_options &o) {
    if (o.solve_mode != HBEM_SOLVE_F64) return o.refine_tolerance;
    return o.refine_tolerance > 0.0 ? std::min(o.refine_tolerance, 1e-13) : 1e-13;
}

// Residual above which a refined system reports refine_stalled: the target,
// and for the FP64-accurate mode 1e-10 (a thousand times below what the FP32
// LU alone leaves, with the FP64 floor out of reach only for systems near
// FP32 singularity).
double stall_threshold(const hbem_solve_options &o) {
    const double tol = refine_tolerance(o);
    return o.solve_mode == HBEM_SOLVE_F64 ? std::max(tol, 1e-10) : tol;
}

// GMRES-based iterative refinement (Carson-Higham GMRES-IR): residuals
// r = b - A x in FP64 with the FP32 operator, correction equations
// M^{-1} A d = M^{-1} r solved by GMRES with M the (low-precision) LU applied
// in FP32, Arnoldi vectors and classical Gram-Schmidt with one
// re-orthogonalisation pass in FP64 on the device; Givens rotations and the
// small least-squares solve on the host.  x64 / r64 hold the iterate and its
// residual on entry and exit; best[b] the relative residual ||r||/||b||
// (inf-norms).  Returns the Arnoldi steps taken per system.
void gmres_refine(hbem_session *s, const hbem::LuWorkspace &w, int B, int vpm, int n, int N,
                  const hbem_solve_options &opt, double tol, int max_restarts,
                  const std::function<void()> &residual, std::vector<double> &nrm,
                  std::vector<double> &best, std::vector<int32_t> &steps,
                  std::vector<char> &stalled) {
    using cd = std::complex<double>;
    hipStream_t st = s->stream;
    const int64_t lda = N, a_stride = (int64_t)N * N;
    const int m = std::max(1, std::min(opt.refine_max_iterations, 20));
    const int64_t vstride = (int64_t)(m + 1) * N;
    s->kv.reserve(sizeof(double) * 2 * (size_t)B * vstride);
    s->kh.reserve(sizeof(double) * 2 * (size_t)B * (m + 1) * 2);
    s->kn.reserve(sizeof(double) * B);
    s->kf.reserve(sizeof(double) * B);
    s->active.reserve(sizeof(int32_t) * B);
    double *V = s->kv.as<double>();
    double *h1 = s->kh.as<double>(), *h2 = h1 + (size_t)B * (m + 1) * 2;
    auto vec = [&](int i) { return V + (size_t)i * N * 2; };  // vector i of system 0
    std::vector<int32_t> active(B);
    std::vector<double> hn(B), fac(B), hbuf((size_t)B * (m + 1) * 2 * 2);
    // M^{-1} src (FP64, stride N) -> basis vector i (FP64)
    auto precond = [&](const double *src, int64_t ld_src, int i) {
        hbem::launch_fill_zero(s->x.as<float>(), (size_t)2 * B * N, st);
        hbem::launch_copy_f64_to_f32(src, s->x.as<float>(), ld_src, N, n, B, st);
        hbem::lu_solve(w, s->x.as<float>(), N, 0, st, vpm, s->trsv_on ? &s->trsv[0] : nullptr);
        hbem::launch_copy_f32_to_f64(s->x.as<float>(), vec(i), N, vstride, n, B, st);
    };
    auto norms_of = [&](int i) {
        hbem::launch_norm2sq_f64(vec(i), vstride, n, B, s->kn.as<double>(), st);
        hip_check(hipMemcpyAsync(hn.data(), s->kn.ptr, sizeof(double) * B, hipMemcpyDeviceToHost,
                                 st),
                  "norm D2H");
        hip_check(hipStreamSynchronize(st), "gmres sync");
    };
    auto scale_vec = [&](int i) {
        hip_check(hipMemcpyAsync(s->kf.ptr, fac.data(), sizeof(double) * B, hipMemcpyHostToDevice,
                                 st),
                  "scale H2D");
        hbem::launch_scale_f64(vec(i), vstride, vec(i), vstride, n, B, s->kf.as<double>(), st);
    };
    for (int restart = 0; restart < max_restarts; ++restart) {
        bool any = false;
        for (int b = 0; b < B; ++b) {
            active[b] = (best[b] > tol && !stalled[b]) ? 1 : 0;
            any |= active[b] != 0;
        }
        if (!any) break;
        hbem::launch_fill_zero_f64(V, (size_t)2 * B * vstride, st);
        precond(s->r64.as<double>(), N, 0);
        norms_of(0);
        std::vector<double> beta(B);
        for (int b = 0; b < B; ++b) {
            beta[b] = std::sqrt(hn[b]);
            fac[b] = beta[b] > 0.0 ? 1.0 / beta[b] : 0.0;
        }
        scale_vec(0);
        std::vector<std::vector<cd>> H(B, std::vector<cd>((size_t)(m + 1) * m, cd(0.0)));
        std::vector<std::vector<cd>> g(B, std::vector<cd>(m + 1, cd(0.0)));
        std::vector<std::vector<double>> cs(B, std::vector<double>(m, 0.0));
        std::vector<std::vector<cd>> sn(B, std::vector<cd>(m, cd(0.0)));
        std::vector<int> kk(B, 0);
        std::vector<char> done(B, 0);
        // inner target: the correction only has to carry the outer residual
        // below the tolerance (with a factor 10 of margin)
        std::vector<double> inner(B);
        for (int b = 0; b < B; ++b) {
            g[b][0] = cd(beta[b], 0.0);
            if (!active[b] || beta[b] == 0.0) done[b] = 1;
            inner[b] = std::max(1e-14, 0.1 * tol / std::max(best[b], 1e-300));
        }
        int j = 0;
        for (; j < m; ++j) {
            // z = M^{-1} A v_j into v_{j+1}, then CGS2 against v_0..v_j
            hbem::launch_gemv_f64(s->A_orig.as<float>(), lda, a_stride, n, B, vec(j), vstride,
                                  s->r64.as<double>() + 0, N, st, vpm);
            precond(s->r64.as<double>(), N, j + 1);
            hbem::launch_dots_f64(V, N, vstride, j + 1, vec(j + 1), vstride, n, B, h1, m + 1, st);
            hbem::launch_combine_f64(V, N, vstride, j + 1, h1, m + 1, -1.0, vec(j + 1), vstride, n,
                                     B, nullptr, st);
            hbem::launch_dots_f64(V, N, vstride, j + 1, vec(j + 1), vstride, n, B, h2, m + 1, st);
            hbem::launch_combine_f64(V, N, vstride, j + 1, h2, m + 1, -1.0, vec(j + 1), vstride, n,
                                     B, nullptr, st);
            hip_check(hipMemcpyAsync(hbuf.data(), h1, sizeof(double) * hbuf.size(),
                                     hipMemcpyDeviceToHost, st),
                      "h D2H");
            norms_of(j + 1);
            bool all_done = true;
            for (int b = 0; b < B; ++b) {
                const double hnext = std::sqrt(hn[b]);
                fac[b] = 0.0;
                if (done[b]) continue;
                std::vector<cd> col(j + 2);
                for (int i = 0; i <= j; ++i) {
                    const size_t o1 = ((size_t)b * (m + 1) + i) * 2;
                    const size_t o2 = o1 + (size_t)B * (m + 1) * 2;
                    col[i] = cd(hbuf[o1] + hbuf[o2], hbuf[o1 + 1] + hbuf[o2 + 1]);
                }
                col[j + 1] = cd(hnext, 0.0);
                // previous rotations, then a new one zeroing col[j+1]
                for (int i = 0; i < j; ++i) {
                    const cd t = cs[b][i] * col[i] + sn[b][i] * col[i + 1];
                    col[i + 1] = -std::conj(sn[b][i]) * col[i] + cs[b][i] * col[i + 1];
                    col[i] = t;
                }
                const double a = std::abs(col[j]), hb = std::abs(col[j + 1]);
                const double den = std::hypot(a, hb);
                if (den == 0.0) {
                    done[b] = 1;
                    continue;
                }
                cs[b][j] = a / den;
                sn[b][j] = (a == 0.0) ? cd(1.0, 0.0) : (col[j] / a) * std::conj(col[j + 1]) / den;
                col[j] = cs[b][j] * col[j] + sn[b][j] * col[j + 1];
                col[j + 1] = cd(0.0);
                g[b][j + 1] = -std::conj(sn[b][j]) * g[b][j];
                g[b][j] = cs[b][j] * g[b][j];
                for (int i = 0; i <= j; ++i) H[b][(size_t)i * m + j] = col[i];
                kk[b] = j + 1;
                steps[b]++;
                if (std::abs(g[b][j + 1]) <= inner[b] * beta[b] || hnext <= 1e-300) {
                    done[b] = 1;
                } else {
                    fac[b] = 1.0 / hnext;
                }
                all_done &= done[b] != 0;
            }
            if (all_done) break;
            scale_vec(j + 1);
        }
        // y = R^{-1} g per system, x += V y
        std::vector<double> ybuf((size_t)B * (m + 1) * 2, 0.0);
        for (int b = 0; b < B; ++b) {
            const int k = kk[b];
            std::vector<cd> y(k);
            for (int i = k - 1; i >= 0; --i) {
                cd acc = g[b][i];
                for (int l = i + 1; l < k; ++l) acc -= H[b][(size_t)i * m + l] * y[l];
                y[i] = acc / H[b][(size_t)i * m + i];
            }
            for (int i = 0; i < k; ++i) {
                ybuf[((size_t)b * (m + 1) + i) * 2] = y[i].real();
                ybuf[((size_t)b * (m + 1) + i) * 2 + 1] = y[i].imag();
            }
        }
        hip_check(hipMemcpyAsync(h1, ybuf.data(), sizeof(double) * ybuf.size(),
                                 hipMemcpyHostToDevice, st),
                  "y H2D");
        hip_check(hipMemcpyAsync(s->active.ptr, active.data(), sizeof(int32_t) * B,
                                 hipMemcpyHostToDevice, st),
                  "active H2D");
        hip_check(hipMemcpyAsync(s->xprev.ptr, s->x64.ptr, sizeof(double) * 2 * (size_t)B * N,
                                 hipMemcpyDeviceToDevice, st),
                  "xprev copy");
        hbem::launch_combine_f64(V, N, vstride, m + 1, h1, m + 1, 1.0, s->x64.as<double>(), N, n,
                                 B, s->active.as<int32_t>(), st);
        residual();
        for (int b = 0; b < B; ++b) {
            if (!active[b]) continue;
            const double rel = nrm[2 * b + 1] > 0.0 ? nrm[2 * b] / nrm[2 * b + 1] : nrm[2 * b];
            if (rel >= best[b]) {
                // no progress: keep the previous iterate and stop refining this
                // system (as the classical loop, main.swift:7051-7058)
                hip_check(hipMemcpyAsync(s->x64.as<double>() + (size_t)b * N * 2,
                                         s->xprev.as<double>() + (size_t)b * N * 2,
                                         sizeof(double) * 2 * (size_t)N, hipMemcpyDeviceToDevice,
                                         st),
                          "revert");
                stalled[b] = 1;
                continue;
            }
            best[b] = rel;
        }
    }
}

// Factor + solve + optional refinement of the B systems resident in s->A
// (column-major, order N, true order n) with the B * S right-hand sides in
// s->rhs (vector v against matrix v / S).  Solutions (FP32, interleaved,
// stride N) end in s->x.  Fills per-matrix diagnostics (refinement figures
// are the worst over the matrix's right-hand sides).
// CHIEF least squares on the resident batch (solveDenseLeastSquaresZgels,
// main.swift:7220-7416, solved differently): with the square system's FP32
// LU and FP64-refined solves, min |A p - b|^2 + |s (C p - d)|^2 has the
// solution p = A^{-1} y with
//   Z = A^{-H} C^H (m refined transposed solves per matrix),
//   y = b - s Z (I + s^2 Z^H Z)^{-1} s (C p0 - d),   p0 = A^{-1} b
// (the normal equations in y = A p, reduced by the push-through identity to
// an m x m Hermitian system).  Writes y into s->rhs; the caller solves
// A p = y again.  chief_res is left for chief_residuals.
namespace {
using cdbl = std::complex<double>;

// Gaussian elimination with partial pivoting, m x m complex (row-major),
// one right-hand side.
void small_solve(std::vector<cdbl> K, std::vector<cdbl> &x, int m) {
    for (int c = 0; c < m; ++c) {
        int piv = c;
        for (int r = c + 1; r < m; ++r)
            if (std::abs(K[(size_t)r * m + c]) > std::abs(K[(size_t)piv * m + c])) piv = r;
        if (piv != c) {
            for (int j = 0; j < m; ++j) std::swap(K[(size_t)c * m + j], K[(size_t)piv * m + j]);
            std::swap(x[c], x[piv]);
        }
        for (int r = c + 1; r < m; ++r) {
            const cdbl f = K[(size_t)r * m + c] / K[(size_t)c * m + c];
            for (int j = c; j < m; ++j) K[(size_t)r * m + j] -= f * K[(size_t)c * m + j];
            x[r] -= f * x[c];
        }
    }
    for (int r = m - 1; r >= 0; --r) {
        cdbl acc = x[r];
        for (int j = r + 1; j < m; ++j) acc -= K[(size_t)r * m + j] * x[j];
        x[r] = acc / K[(size_t)r * m + r];
    }
}
}  // namespace

void chief_solve(hbem_session *s, const hbem::LuWorkspace &w, int B, int S, int n, int N,
                 const hbem_solve_options &opt, std::vector<double> *chief_res) {
    hipStream_t st = s->stream;
    const int m = opt.n_chief, V = B * S, BM = B * m;
    const int64_t lda = N, a_stride = (int64_t)N * N;
    const double wgt = opt.chief_weight > 0.0 ? opt.chief_weight : 1.0;
    std::vector<float> Ch((size_t)2 * BM * n), dh((size_t)2 * V * m);
    std::vector<double> ainf(B);
    s->chief_ainf.reserve(sizeof(double) * B);
    hbem::launch_row_abs_sum(s->A_orig.as<float>(), lda, a_stride, n, B, s->chief_ainf.as<double>(),
                             st);
    hip_check(hipMemcpyAsync(Ch.data(), s->chief_C.ptr, sizeof(float) * Ch.size(),
                             hipMemcpyDeviceToHost, st), "C D2H");
    hip_check(hipMemcpyAsync(dh.data(), s->chief_d.ptr, sizeof(float) * dh.size(),
                             hipMemcpyDeviceToHost, st), "d D2H");
    hip_check(hipMemcpyAsync(ainf.data(), s->chief_ainf.ptr, sizeof(double) * B,
                             hipMemcpyDeviceToHost, st), "norm D2H");
    hip_check(hipStreamSynchronize(st), "chief sync");
    // Z = A^{-H} C^H: right-hand sides conj(C) rows, float for the LU and
    // double for the refinement target
    std::vector<float> zr((size_t)2 * BM * N, 0.0f);
    std::vector<double> ze((size_t)2 * BM * N, 0.0);
    for (int bm = 0; bm < BM; ++bm)
        for (int j = 0; j < n; ++j) {
            const float re = Ch[((size_t)bm * n + j) * 2], im = Ch[((size_t)bm * n + j) * 2 + 1];
            zr[((size_t)bm * N + j) * 2] = re;
            zr[((size_t)bm * N + j) * 2 + 1] = -im;
            ze[((size_t)bm * N + j) * 2] = re;
            ze[((size_t)bm * N + j) * 2 + 1] = -im;
        }
    upload(s->chief_z, zr, st);
    upload(s->chief_e, ze, st);
    s->chief_z64.reserve(sizeof(double) * ze.size());
    s->chief_r.reserve(sizeof(double) * ze.size());
    s->norms.reserve(sizeof(double) * 2 * std::max(BM, V));
    hbem::lu_solve(w, s->chief_z.as<float>(), N, 2, st, m);
    hbem::launch_fill_zero_f64(s->chief_z64.as<double>(), ze.size(), st);
    hbem::launch_axpy_f32_to_f64(s->chief_z.as<float>(), s->chief_z64.as<double>(), N, N, n, BM,
                                 nullptr, st);
    for (int it = 0; it < 3; ++it) {
        hbem::launch_residual_h_f64(s->A_orig.as<float>(), lda, a_stride, n, BM,
                                    s->chief_e.as<double>(), s->chief_z64.as<double>(), N,
                                    s->chief_r.as<double>(), s->norms.as<double>(), st, m);
        hbem::launch_fill_zero(s->chief_z.as<float>(), zr.size(), st);
        hbem::launch_copy_f64_to_f32(s->chief_r.as<double>(), s->chief_z.as<float>(), N, N, n, BM,
                                     st);
        hbem::lu_solve(w, s->chief_z.as<float>(), N, 2, st, m);
        hbem::launch_axpy_f32_to_f64(s->chief_z.as<float>(), s->chief_z64.as<double>(), N, N, n,
                                     BM, nullptr, st);
    }
    std::vector<double> Z(ze.size()), x0((size_t)2 * V * N);
    std::vector<float> bh((size_t)2 * V * N);
    hip_check(hipMemcpyAsync(Z.data(), s->chief_z64.ptr, sizeof(double) * Z.size(),
                             hipMemcpyDeviceToHost, st), "Z D2H");
    hip_check(hipMemcpyAsync(x0.data(), s->x64.ptr, sizeof(double) * x0.size(),
                             hipMemcpyDeviceToHost, st), "x D2H");
    hip_check(hipMemcpyAsync(bh.data(), s->rhs.ptr, sizeof(float) * bh.size(),
                             hipMemcpyDeviceToHost, st), "rhs D2H");
    hip_check(hipStreamSynchronize(st), "chief sync");
    s->chief_b = bh;
    s->chief_scale.assign(B, 1.0);
    std::vector<float> y(bh);
    for (int b = 0; b < B; ++b) {
        double cinf = 0.0;
        for (int i = 0; i < m; ++i) {
            double row = 0.0;
            for (int j = 0; j < n; ++j) {
                const size_t o = (((size_t)b * m + i) * n + j) * 2;
                row += std::hypot((double)Ch[o], (double)Ch[o + 1]);
            }
            cinf = std::max(cinf, row);
        }
        const double sc = wgt * (cinf > 0.0 ? ainf[b] / cinf : 1.0);
        s->chief_scale[b] = sc;
        auto zat = [&](int i, int j) {
            const size_t o = (((size_t)b * m + i) * N + j) * 2;
            return cdbl(Z[o], Z[o + 1]);
        };
        // K = I + s^2 Z^H Z
        std::vector<cdbl> K((size_t)m * m);
        for (int i = 0; i < m; ++i)
            for (int k = 0; k < m; ++k) {
                cdbl acc = 0.0;
                for (int j = 0; j < n; ++j) acc += std::conj(zat(i, j)) * zat(k, j);
                K[(size_t)i * m + k] = sc * sc * acc + (i == k ? 1.0 : 0.0);
            }
        for (int src = 0; src < S; ++src) {
            const int v = b * S + src;
            std::vector<cdbl> r(m);
            for (int i = 0; i < m; ++i) {
                cdbl acc = 0.0;
                for (int j = 0; j < n; ++j) {
                    const size_t oc = (((size_t)b * m + i) * n + j) * 2;
                    const size_t ox = ((size_t)v * N + j) * 2;
                    acc += cdbl(Ch[oc], Ch[oc + 1]) * cdbl(x0[ox], x0[ox + 1]);
                }
                const size_t od = ((size_t)v * m + i) * 2;
                r[i] = sc * (acc - cdbl(dh[od], dh[od + 1]));
            }
            small_solve(K, r, m);
            for (int j = 0; j < n; ++j) {
                cdbl acc = 0.0;
                for (int i = 0; i < m; ++i) acc += zat(i, j) * r[i];
                const size_t o = ((size_t)v * N + j) * 2;
                y[o] = (float)((double)bh[o] - sc * acc.real());
                y[o + 1] = (float)((double)bh[o + 1] - sc * acc.imag());
            }
        }
    }
    hip_check(hipMemcpyAsync(s->rhs.ptr, y.data(), sizeof(float) * y.size(), hipMemcpyHostToDevice,
                             st), "rhs H2D");
    hip_check(hipStreamSynchronize(st), "chief sync");
    (void)chief_res;
}

// |s (C p - d)|_2 / |b|_2 per vector with the final solution (s->x64).
void chief_residuals(hbem_session *s, int B, int S, int n, int N, const hbem_solve_options &opt,
                     std::vector<double> *chief_res) {
    hipStream_t st = s->stream;
    const int m = opt.n_chief, V = B * S;
    std::vector<float> Ch((size_t)2 * B * m * n), dh((size_t)2 * V * m);
    std::vector<double> x((size_t)2 * V * N);
    hip_check(hipMemcpyAsync(Ch.data(), s->chief_C.ptr, sizeof(float) * Ch.size(),
                             hipMemcpyDeviceToHost, st), "C D2H");
    hip_check(hipMemcpyAsync(dh.data(), s->chief_d.ptr, sizeof(float) * dh.size(),
                             hipMemcpyDeviceToHost, st), "d D2H");
    hip_check(hipMemcpyAsync(x.data(), s->x64.ptr, sizeof(double) * x.size(),
                             hipMemcpyDeviceToHost, st), "x D2H");
    hip_check(hipStreamSynchronize(st), "chief sync");
    for (int v = 0; v < V; ++v) {
        const int b = v / S;
        double num = 0.0, den = 0.0;
        for (int i = 0; i < m; ++i) {
            cdbl acc = 0.0;
            for (int j = 0; j < n; ++j) {
                const size_t oc = (((size_t)b * m + i) * n + j) * 2;
                const size_t ox = ((size_t)v * N + j) * 2;
                acc += cdbl(Ch[oc], Ch[oc + 1]) * cdbl(x[ox], x[ox + 1]);
            }
            const size_t od = ((size_t)v * m + i) * 2;
            num += std::norm(s->chief_scale[b] * (acc - cdbl(dh[od], dh[od + 1])));
        }
        for (int j = 0; j < n; ++j) {
            const size_t o = ((size_t)v * N + j) * 2;
            den += (double)s->chief_b[o] * s->chief_b[o] + (double)s->chief_b[o + 1] * s->chief_b[o + 1];
        }
        (*chief_res)[v] = den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
    }
}

void solve_resident(hbem_session *s, int B, int S, int n, int N, const hbem_solve_options &opt,
                    std::vector<hbem_case_diag> *diag, double *solve_seconds,
                    const std::function<void()> &after_solve = nullptr) {
    hipStream_t st = s->stream;
    const int64_t lda = N, a_stride = (int64_t)N * N;
    const int V = B * S;
    // CHIEF solves in the FP64-refined mode, as the reference's least squares
    // runs in FP64 (sweep.py:123)
    const bool refine = opt.solve_mode != HBEM_SOLVE_F32 || opt.n_chief > 0;
    const bool gmres = opt.solve_mode == HBEM_SOLVE_MIXED_F16 ||
                       opt.solve_mode == HBEM_SOLVE_MIXED_BF16 ||
                       opt.solve_mode == HBEM_SOLVE_F64;
    const int precision = opt.solve_mode == HBEM_SOLVE_MIXED_F16    ? 1
                          : opt.solve_mode == HBEM_SOLVE_MIXED_BF16 ? 2
                                                                    : 0;
    Timer t;
    t.start(st);
    const bool keep_a = refine || s->adj_keep;
    s->anorm.reserve(sizeof(double) * B);
    s->amax.reserve(sizeof(double) * B);
    s->lu_scale.reserve(sizeof(float) * B);
    if (keep_a) s->A_orig.reserve(sizeof(float) * 2 * (size_t)B * a_stride);
    s->ipiv.reserve(sizeof(int32_t) * (size_t)B * N);
    s->info.reserve(sizeof(int32_t) * B);
    s->gl.reserve(sizeof(int32_t) * (size_t)B * hbem::lu_gather_ints(N));
    s->lu_scr_i.reserve(sizeof(int32_t) * (size_t)B * hbem::kLuScratchI);
    s->lu_scr_f.reserve(sizeof(float) * (size_t)B * hbem::kLuScratchF);
    s->x.reserve(sizeof(float) * 2 * (size_t)V * N);
    hbem::LuWorkspace w{s->A.as<float>(), lda, a_stride, N, n, B, s->ipiv.as<int32_t>(),
                        s->info.as<int32_t>(), s->gl.as<int32_t>(), s->lu_scr_i.as<int32_t>(),
                        s->lu_scr_f.as<float>()};
    if (precision != 0 && N > hbem::kLuOuter) {
        s->lu_pack.reserve(sizeof(uint16_t) * (size_t)B * hbem::lu_pack_u16(N));
        w.pack = s->lu_pack.as<uint16_t>();
    }
    // multi-workgroup triangular solves: row maps built by the factorization,
    // one solve context per stream
    s->trsv_on = N <= hbem::kTrsvMaxOrder;
    if (s->trsv_on) {
        s->lu_perm.reserve(sizeof(int32_t) * (size_t)B * N);
        s->lu_pi.reserve(sizeof(int32_t) * (size_t)B * hbem::lu_solve_blocks(N) * N);
        s->lu_dinv.reserve(sizeof(float) * (size_t)B * hbem::lu_dinv_floats(N));
        s->lu_lf.reserve(sizeof(float) * 2 * (size_t)B * a_stride);
        w.lf = s->lu_lf.as<float>();
        w.perm = s->lu_perm.as<int32_t>();
        w.pi = s->lu_pi.as<int32_t>();
        w.dinv = s->lu_dinv.as<float>();
        s->trsv[0] = trsv_sync(s, 0, V, N, st);
        s->trsv[1] = trsv_sync(s, 1, B, N, st);
    }
    // The rcond iteration needs only the LU factors: it runs on the aux
    // stream from its own host thread, from the end of the factorization on,
    // while this thread solves, refines, solves the CHIEF rows and enqueues
    // the field (the factors are read-only from there on; the two sides use
    // disjoint buffers).
    const bool want_rcond = opt.estimate_rcond != 0;
    if (want_rcond) {
        if (!s->aux_stream)
            hip_check(hipStreamCreateWithFlags(&s->aux_stream, hipStreamNonBlocking),
                      "hipStreamCreate");
        if (!s->ev_lu)
            hip_check(hipEventCreateWithFlags(&s->ev_lu, hipEventDisableTiming), "hipEventCreate");
        reserve_rcond(s, B, N);
    }
    // padding, norms, FP16 operand scale, LU, first solve: kernels and
    // device copies only (no host synchronisation, no memset nodes);
    // mark_lu records the end of the factorization for the rcond stream
    auto enqueue = [&](bool mark_lu) {
        hbem::launch_pad_identity(s->A.as<float>(), lda, a_stride, n, N, B, st);
        hbem::launch_norm1(s->A.as<float>(), lda, a_stride, n, B, s->anorm.as<double>(),
                           s->amax.as<double>(), st);
        hbem::launch_pow2_scale(s->amax.as<double>(), s->lu_scale.as<float>(), B, st);
        if (keep_a)
            hip_check(hipMemcpyAsync(s->A_orig.ptr, s->A.ptr,
                                     sizeof(float) * 2 * (size_t)B * a_stride,
                                     hipMemcpyDeviceToDevice, st),
                      "A copy");
        hbem::lu_factor(w, precision, s->lu_scale.as<float>(), st);
        if (mark_lu) hip_check(hipEventRecord(s->ev_lu, st), "hipEventRecord");
        hip_check(hipMemcpyAsync(s->x.ptr, s->rhs.ptr, sizeof(float) * 2 * (size_t)V * N,
                                 hipMemcpyDeviceToDevice, st),
                  "rhs copy");
        hbem::lu_solve(w, s->x.as<float>(), N, opt.solve_trans, st, S,
                       s->trsv_on ? &s->trsv[0] : nullptr);
    };
    if (opt.lu_graph) {
        // One hipGraph per batch shape and buffer set, replayed while they
        // stay the same (sweeps of equal-size chunks, repeated sweeps).
        const std::vector<uintptr_t> key = {
            (uintptr_t)B, (uintptr_t)N, (uintptr_t)n, (uintptr_t)S, (uintptr_t)precision,
            (uintptr_t)keep_a, (uintptr_t)s->A.ptr, (uintptr_t)s->A_orig.ptr,
            (uintptr_t)s->rhs.ptr, (uintptr_t)s->x.ptr, (uintptr_t)s->ipiv.ptr,
            (uintptr_t)s->info.ptr, (uintptr_t)s->gl.ptr, (uintptr_t)s->lu_scr_i.ptr,
            (uintptr_t)s->lu_scr_f.ptr, (uintptr_t)s->anorm.ptr, (uintptr_t)s->amax.ptr,
            (uintptr_t)s->lu_scale.ptr, (uintptr_t)w.pack, (uintptr_t)w.perm, (uintptr_t)w.pi, (uintptr_t)w.dinv, (uintptr_t)w.lf,
            (uintptr_t)s->ts_ctx[0].ptr, (uintptr_t)s->ts_y[0].ptr, (uintptr_t)s->ts_z[0].ptr};
        if (!s->lu_graph || s->lu_graph_key != key) {
            if (s->lu_graph) {
                (void)hipGraphExecDestroy(s->lu_graph);
                s->lu_graph = nullptr;
            }
            hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture begin");
            hipGraph_t graph = nullptr;
            try {
                enqueue(false);
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            hip_check(hipStreamEndCapture(st, &graph), "capture end");
            const hipError_t e = hipGraphInstantiate(&s->lu_graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(e, "graph instantiate");
            s->lu_graph_key = key;
        }
        hip_check(hipGraphLaunch(s->lu_graph, st), "graph launch");
        if (want_rcond) hip_check(hipEventRecord(s->ev_lu, st), "hipEventRecord");
    } else {
        enqueue(want_rcond);
    }
    std::vector<double> est;
    double rcond_s = 0.0;
    std::exception_ptr rc_error;
    std::thread rc_thread;
    if (want_rcond) {
        rc_thread = std::thread([&] {
            try {
                hip_check(hipSetDevice(s->device), "hipSetDevice");
                hip_check(hipStreamWaitEvent(s->aux_stream, s->ev_lu, 0), "hipStreamWaitEvent");
                Timer tr;
                tr.start(s->aux_stream);
                estimate_rcond(s, w, &est, s->aux_stream);
                tr.stop(s->aux_stream);
                rcond_s = tr.seconds();
            } catch (...) {
                rc_error = std::current_exception();
            }
        });
    }
    struct Join {
        std::thread &t;
        ~Join() {
            if (t.joinable()) t.join();
        }
    } join_rcond{rc_thread};
    std::vector<int32_t> info(B);
    hip_check(hipMemcpyAsync(info.data(), s->info.ptr, sizeof(int32_t) * B, hipMemcpyDeviceToHost,
                             st),
              "info D2H");
    std::vector<double> anorm(B);
    hip_check(hipMemcpyAsync(anorm.data(), s->anorm.ptr, sizeof(double) * B,
                             hipMemcpyDeviceToHost, st),
              "anorm D2H");
    hip_check(hipStreamSynchronize(st), "solve sync");
    std::vector<int32_t> iters(V, 0);
    std::vector<double> resid(V, -1.0);
    std::vector<char> stalled(V, 0);
    auto refine_all = [&]() {
        if (!refine || opt.refine_max_iterations <= 0) return;
        s->x64.reserve(sizeof(double) * 2 * (size_t)V * N);
        s->r64.reserve(sizeof(double) * 2 * (size_t)V * N);
        s->xprev.reserve(sizeof(double) * 2 * (size_t)V * N);
        s->norms.reserve(sizeof(double) * 2 * V);
        s->active.reserve(sizeof(int32_t) * V);
        hbem::launch_fill_zero_f64(s->x64.as<double>(), (size_t)2 * V * N, st);
        hbem::launch_copy_f32_to_f64(s->x.as<float>(), s->x64.as<double>(), N, N, n, V, st);
        std::vector<double> best(V, 0.0), nrm(2 * V);
        std::vector<int32_t> active(V, 1);
        std::fill(stalled.begin(), stalled.end(), 0);
        auto residual = [&]() {
            hbem::launch_residual_f64(s->A_orig.as<float>(), lda, a_stride, n, V,
                                      s->rhs.as<float>(), N, s->x64.as<double>(), N,
                                      s->r64.as<double>(), s->norms.as<double>(), st, S);
            hip_check(hipMemcpyAsync(nrm.data(), s->norms.ptr, sizeof(double) * 2 * V,
                                     hipMemcpyDeviceToHost, st),
                      "norms D2H");
            hip_check(hipStreamSynchronize(st), "residual sync");
        };
        residual();
        const double tol = refine_tolerance(opt);
        for (int v = 0; v < V; ++v) {
            best[v] = nrm[2 * v + 1] > 0.0 ? nrm[2 * v] / nrm[2 * v + 1] : nrm[2 * v];
            if (best[v] <= tol) active[v] = 0;
        }
        if (gmres) {
            // mixed precision and the FP64-accurate mode: GMRES-IR with the
            // FP32-stored LU as preconditioner
            gmres_refine(s, w, V, S, n, N, opt, tol, opt.solve_mode == HBEM_SOLVE_F64 ? 8 : 4,
                         residual, nrm, best, iters, stalled);
        } else {
            // FP32 LU: classical refinement as the reference's
            // refineDenseSolveSolution (main.swift:6964-7070)
            for (int it = 0; it < opt.refine_max_iterations; ++it) {
                bool any = false;
                for (int v = 0; v < V; ++v) any |= active[v] != 0;
                if (!any) break;
                hip_check(hipMemcpyAsync(s->active.ptr, active.data(), sizeof(int32_t) * V,
                                         hipMemcpyHostToDevice, st),
                          "active H2D");
                hip_check(hipMemcpyAsync(s->xprev.ptr, s->x64.ptr,
                                         sizeof(double) * 2 * (size_t)V * N,
                                         hipMemcpyDeviceToDevice, st),
                          "xprev copy");
                hbem::launch_fill_zero(s->x.as<float>(), (size_t)2 * V * N, st);
                hbem::launch_copy_f64_to_f32(s->r64.as<double>(), s->x.as<float>(), N, N, n, V,
                                             st);
                hbem::lu_solve(w, s->x.as<float>(), N, 0, st, S,
                               s->trsv_on ? &s->trsv[0] : nullptr);
                hbem::launch_axpy_f32_to_f64(s->x.as<float>(), s->x64.as<double>(), N, N, n, V,
                                             s->active.as<int32_t>(), st);
                residual();
                for (int v = 0; v < V; ++v) {
                    if (!active[v]) continue;
                    const double rel =
                        nrm[2 * v + 1] > 0.0 ? nrm[2 * v] / nrm[2 * v + 1] : nrm[2 * v];
                    if (rel >= best[v]) {
                        // stalled: keep the previous iterate (main.swift:7051-7058)
                        hip_check(hipMemcpyAsync(s->x64.as<double>() + (size_t)v * N * 2,
                                                 s->xprev.as<double>() + (size_t)v * N * 2,
                                                 sizeof(double) * 2 * (size_t)N,
                                                 hipMemcpyDeviceToDevice, st),
                                  "revert");
                        stalled[v] = 1;
                        active[v] = 0;
                        continue;
                    }
                    best[v] = rel;
                    iters[v]++;
                    if (rel <= tol) active[v] = 0;
                }
            }
        }
        hbem::launch_fill_zero(s->x.as<float>(), (size_t)2 * V * N, st);
        hbem::launch_copy_f64_to_f32(s->x64.as<double>(), s->x.as<float>(), N, N, n, V, st);
        const double stall_at = stall_threshold(opt);
        for (int v = 0; v < V; ++v) {
            resid[v] = best[v];
            // stopped above the target (no progress, or out of iterations)
            stalled[v] = best[v] > stall_at ? 1 : 0;
        }
    };
    refine_all();
    std::vector<double> chief_res(V, -1.0);
    if (opt.n_chief > 0) {
        chief_solve(s, w, B, S, n, N, opt, &chief_res);
        hip_check(hipMemcpyAsync(s->x.ptr, s->rhs.ptr, sizeof(float) * 2 * (size_t)V * N,
                                 hipMemcpyDeviceToDevice, st),
                  "rhs copy");
        hbem::lu_solve(w, s->x.as<float>(), N, 0, st, S, s->trsv_on ? &s->trsv[0] : nullptr);
        refine_all();
        chief_residuals(s, B, S, n, N, opt, &chief_res);
    }
    t.stop(st);
    *solve_seconds = t.seconds();
    if (after_solve) after_solve();
    if (rc_thread.joinable()) rc_thread.join();
    if (rc_error) std::rethrow_exception(rc_error);
    std::vector<double> rcond(B, -1.0);
    if (want_rcond)
        for (int b = 0; b < B; ++b)
            rcond[b] = (anorm[b] > 0.0 && est[b] > 0.0) ? (1.0 / est[b]) / anorm[b] : 0.0;
    diag->resize(B);
    for (int b = 0; b < B; ++b) {
        hbem_case_diag &d = (*diag)[b];
        d.lapack_info = info[b];
        d.refine_iterations = 0;
        d.refine_residual_rel = -1.0;
        d.refine_stalled = 0;
        d.rcond_s = rcond_s / B;
        for (int src = 0; src < S; ++src) {
            d.refine_stalled |= stalled[(size_t)b * S + src];
            d.refine_iterations = std::max(d.refine_iterations, iters[(size_t)b * S + src]);
            d.refine_residual_rel = std::max(d.refine_residual_rel, resid[(size_t)b * S + src]);
        }
        d.dense_solve_condition_1norm = anorm[b];
        d.dense_solve_rcond = rcond[b];
        d.chief_residual_rel = -1.0;
        for (int src = 0; src < S; ++src)
            d.chief_residual_rel = std::max(d.chief_residual_rel, chief_res[(size_t)b * S + src]);
    }
}

int batch_size_for(const hbem_session *s, int F, int N, const hbem_solve_options &opt) {
    if (s->adj_keep) return F;
    if (opt.batch_frequencies > 0) return std::min(F, opt.batch_frequencies);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)16 << 30;
    const bool refine = opt.solve_mode != HBEM_SOLVE_F32;
    // A, the refinement's copy of A, L in final row order for the
    // multi-workgroup solves, the 16-bit LU operands
    const double mats = 1.0 + (refine ? 1.0 : 0.0) + (N <= hbem::kTrsvMaxOrder ? 1.0 : 0.0);
    const double per = (double)N * N * 8.0 * mats + (double)N * 64.0 +
                       (refine ? 2.0 * (double)hbem::lu_pack_u16(N) : 0.0);
    const double budget = std::min(0.6 * (double)free_b, 192.0 * (1ull << 30));
    int b = (int)std::max(1.0, std::floor(budget / per));
    return std::max(1, std::min({F, b, 64}));
}


void check_options(const hbem_solve_options &o) {
    if (o.solve_mode < HBEM_SOLVE_F32 || o.solve_mode > HBEM_SOLVE_F64)
        fail(HBEM_EINVAL, "solve_mode out of range");
    if (o.solve_mode == HBEM_SOLVE_F64 && o.refine_max_iterations <= 0)
        fail(HBEM_EINVAL, "the FP64-accurate solve needs refine_max_iterations > 0");
    if (o.solve_trans != 0 && o.solve_trans != 2) fail(HBEM_EINVAL, "solve_trans must be 0 or 2");
    if (o.solve_trans == 2 && (o.solve_mode != HBEM_SOLVE_F32 || o.n_chief > 0))
        fail(HBEM_EINVAL, "conjugate-transpose solves take solve_mode FP32 without CHIEF rows");
    if (o.refine_max_iterations < 0) fail(HBEM_EINVAL, "refine_max_iterations must be >= 0");
    if (!(o.refine_tolerance >= 0.0)) fail(HBEM_EINVAL, "refine_tolerance must be >= 0");
    if (o.n_tags < 0 || (o.n_tags > 0 && !o.tags)) fail(HBEM_EINVAL, "tags invalid");
    if (o.n_sources < 0 || o.n_sources > 4096) fail(HBEM_EINVAL, "n_sources out of range");
    if (o.n_chief < 0 || o.n_chief > 4096) fail(HBEM_EINVAL, "n_chief out of range");
    if (o.precise_geometry != 0 && o.precise_geometry != 1)
        fail(HBEM_EINVAL, "precise_geometry must be 0 or 1");
    if (o.n_chief > 0) {
        if (!o.chief_points) fail(HBEM_EINVAL, "chief_points must be given with n_chief > 0");
        if (!(std::isfinite(o.chief_weight) && o.chief_weight >= 0.0))
            fail(HBEM_EINVAL, "chief_weight must be finite and positive");
        if (o.refine_max_iterations <= 0)
            fail(HBEM_EINVAL, "CHIEF needs refine_max_iterations > 0 (FP64-refined solves)");
    }
}

// All-device batch pipeline.  Inputs: k_real/field_k [F], k_imag [F] or
// null, q [F][S][T] interleaved complex, beta [F][T] interleaved complex or
// null, obs [3][n_obs].  Outputs: field [F][S][n_obs] interleaved complex
// (may be null when n_obs == 0), p [F][S][n] interleaved complex or null,
// tag_avg [F][S][1 + n_tags][2] double (impedance tag first).  Frequencies
// are processed in resident chunks; the S right-hand sides of a frequency
// share its matrix and LU.
void solve_core(hbem_session *s, int F, const float *kre_d, const float *kim_d,
                const float *fk_d, const float *q_d, const float *beta_d, int n_obs,
                const float *obs_d, const hbem_solve_options &opts, float *field_d, float *p_d,
                double *tag_d, std::vector<hbem_case_diag> *diag_out) {
    hipStream_t st = s->stream;
    const int n = s->plan.n_dofs;
    const int T = s->plan.n_triangles;
    const int S = std::max(1, opts.n_sources);
    const int N = (int)round_up(n, hbem::kLuPad);
    const int64_t lda = N, a_stride = (int64_t)N * N;
    const int Bmax = batch_size_for(s, F, N, opts);
    std::vector<int32_t> tag_list;
    tag_list.push_back(opts.impedance_tag);
    for (int i = 0; i < opts.n_tags; ++i) tag_list.push_back(opts.tags[i]);
    const int ntag = (int)tag_list.size();
    upload(s->tags, tag_list, st);
    diag_out->assign(F, hbem_case_diag{});
    if (n_obs > 0) hbem::launch_fill_zero(field_d, (size_t)2 * F * S * n_obs, st);
    for (int f0 = 0; f0 < F; f0 += Bmax) {
        const int B = std::min(Bmax, F - f0), V = B * S;
        const float *qc = q_d + (size_t)f0 * S * T * 2;
        const float *bc = beta_d ? beta_d + (size_t)f0 * T * 2 : nullptr;
        Timer ta, treg, tsing, tnear, tf;
        ta.start(st);
        assemble_device(s, B, S, kre_d + f0, kim_d ? kim_d + f0 : nullptr, qc, bc, lda, a_stride,
                        N, &treg, &tsing, &tnear, opts.precise_geometry != 0 || s->adj_keep);
        if (opts.n_chief > 0) {
            const int m = opts.n_chief;
            s->chief_C.reserve(sizeof(float) * 2 * (size_t)B * m * n);
            s->chief_d.reserve(sizeof(float) * 2 * (size_t)V * m);
            hbem::launch_fill_zero(s->chief_C.as<float>(), (size_t)2 * B * m * n, st);
            hbem::launch_fill_zero(s->chief_d.as<float>(), (size_t)2 * V * m, st);
            hbem::launch_chief_rows(s->dg, B, m, s->chief_pts_d, kre_d + f0,
                                    kim_d ? kim_d + f0 : nullptr, S, qc, bc, s->chief_C.as<float>(),
                                    s->chief_d.as<float>(), st);
        }
        ta.stop(st);
        std::vector<hbem_case_diag> dg;
        double solve_s = 0.0;
        // field and reductions: enqueued while the rcond estimate may still run
        auto field_stage = [&]() {
            tf.start(st);
            const float *qf = qc;
            if (bc) {
                s->qt.reserve(sizeof(float) * 2 * (size_t)V * T);
                hbem::launch_robin_neumann(s->dg, V, S, kre_d + f0, kim_d ? kim_d + f0 : nullptr,
                                           qc, bc, s->x.as<float>(), N, s->qt.as<float>(), st);
                qf = s->qt.as<float>();
            }
            if (n_obs > 0)
                hbem::launch_field(s->dg, V, fk_d + f0, S, s->x.as<float>(), N, qf, n_obs, obs_d,
                                   field_d + (size_t)f0 * S * n_obs * 2, st);
            hbem::launch_tag_average(s->dg, V, s->x.as<float>(), N, ntag, s->tags.as<int32_t>(),
                                     tag_d + (size_t)f0 * S * ntag * 2, st);
            if (p_d)
                hip_check(hipMemcpy2DAsync(p_d + (size_t)f0 * S * n * 2, sizeof(float) * 2 * n,
                                           s->x.ptr, sizeof(float) * 2 * N, sizeof(float) * 2 * n,
                                           V, hipMemcpyDeviceToDevice, st),
                          "p copy");
            hip_check(hipGetLastError(), "field launch");
            tf.stop(st);
        };
        solve_resident(s, B, S, n, N, opts, &dg, &solve_s, field_stage);
        const double asm_s = ta.seconds(), reg_s = treg.seconds(), sing_s = tsing.seconds();
        const double near_s = tnear.seconds();
        const double field_s = tf.seconds();
        for (int b = 0; b < B; ++b) {
            hbem_case_diag &d = (*diag_out)[f0 + b];
            d = dg[b];
            d.assembly_s = asm_s / B;
            d.regular_assembly_s = reg_s / B;
            d.singular_assembly_s = sing_s / B;
            d.near_assembly_s = near_s / B;
            d.near_pairs = s->near_pairs;
            d.dense_solve_s = solve_s / B;
            d.field_s = field_s / B;
            d.batch_index = f0 / Bmax;
            d.batch_size = B;
        }
    }
}

}  // namespace

extern "C" {

int32_t hbem_solve_batch(hbem_session *s, int32_t F, const float *k_real, const float *k_imag,
                         const float *field_k, const float *q_re, const float *q_im, int32_t n_obs,
                         const float *obs_3xn, const hbem_solve_options *opts, float *field_re,
                         float *field_im, float *p_re, float *p_im, double *impedance,
                         double *pavg, hbem_case_diag *diag) {
    return guarded([&] {
        if (!s || !opts) fail(HBEM_EINVAL, "hbem_solve_batch: null argument");
        if (F <= 0) fail(HBEM_EINVAL, "hbem_solve_batch: n_freq must be positive");
        if (!k_real || !field_k || !q_re || !q_im)
            fail(HBEM_EINVAL, "hbem_solve_batch: k_real, field_k and q must be non-null");
        if (n_obs < 0 || (n_obs > 0 && (!obs_3xn || !field_re || !field_im)))
            fail(HBEM_EINVAL, "hbem_solve_batch: observation arrays invalid");
        check_options(*opts);
        if (opts->solve_trans != 0) fail(HBEM_EINVAL, "hbem_solve_batch: solve_trans must be 0");
        for (int f = 0; f < F; ++f) {
            if (!std::isfinite(k_real[f]) || !std::isfinite(field_k[f]))
                fail(HBEM_EINVAL, "hbem_solve_batch: wavenumbers must be finite");
            if (k_imag && (!std::isfinite(k_imag[f]) || k_imag[f] < 0.0f))
                fail(HBEM_EINVAL, "hbem_solve_batch: k_imag must be finite and >= 0");
        }
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        set_device(s);
        hipStream_t st = s->stream;
        const int n = s->plan.n_dofs;
        const int T = s->plan.n_triangles;
        const int S = std::max(1, opts->n_sources);
        const bool complex_k = any_nonzero(k_imag, (size_t)F);
        std::vector<float> kr(k_real, k_real + F), fk(field_k, field_k + F);
        upload(s->io_k, kr, st);
        upload(s->io_fk, fk, st);
        if (complex_k) {
            std::vector<float> ki(k_imag, k_imag + F);
            upload(s->io_kim, ki, st);
        }
        upload(s->io_q, interleave(q_re, q_im, (size_t)F * S * T), st);
        const bool robin = any_nonzero(opts->robin_beta, (size_t)2 * F * T);
        if (robin)
            upload(s->io_beta, std::vector<float>(opts->robin_beta, opts->robin_beta + (size_t)2 * F * T),
                   st);
        if (n_obs > 0) upload(s->io_obs, std::vector<float>(obs_3xn, obs_3xn + 3 * (size_t)n_obs), st);
        const int ntag = 1 + opts->n_tags;
        const size_t FS = (size_t)F * S;
        s->io_field.reserve(sizeof(float) * 2 * std::max<size_t>(1, FS * n_obs));
        s->io_tag.reserve(sizeof(double) * 2 * FS * ntag);
        const bool want_p = p_re && p_im;
        if (want_p) s->io_p.reserve(sizeof(float) * 2 * FS * n);
        std::vector<hbem_case_diag> dg;
        if (opts->n_chief > 0) {
            upload(s->chief_pts, std::vector<float>(opts->chief_points,
                                                    opts->chief_points + 3 * (size_t)opts->n_chief),
                   st);
            s->chief_pts_d = s->chief_pts.as<float>();
        }
        solve_core(s, F, s->io_k.as<float>(), complex_k ? s->io_kim.as<float>() : nullptr,
                   s->io_fk.as<float>(), s->io_q.as<float>(),
                   robin ? s->io_beta.as<float>() : nullptr, n_obs, s->io_obs.as<float>(), *opts,
                   s->io_field.as<float>(), want_p ? s->io_p.as<float>() : nullptr,
                   s->io_tag.as<double>(), &dg);
        std::vector<float> fh((size_t)2 * FS * std::max(n_obs, 0));
        if (n_obs > 0)
            hip_check(hipMemcpyAsync(fh.data(), s->io_field.ptr, sizeof(float) * fh.size(),
                                     hipMemcpyDeviceToHost, st),
                      "field D2H");
        std::vector<double> th((size_t)2 * FS * ntag);
        hip_check(hipMemcpyAsync(th.data(), s->io_tag.ptr, sizeof(double) * th.size(),
                                 hipMemcpyDeviceToHost, st),
                  "tags D2H");
        std::vector<float> ph;
        if (want_p) {
            ph.resize((size_t)2 * FS * n);
            hip_check(hipMemcpyAsync(ph.data(), s->io_p.ptr, sizeof(float) * ph.size(),
                                     hipMemcpyDeviceToHost, st),
                      "p D2H");
        }
        hip_check(hipStreamSynchronize(st), "batch sync");
        for (int f = 0; f < F; ++f) {
            if (dg[f].lapack_info != 0 && !diag)
                fail(HBEM_ESINGULAR, "LU found an exactly zero pivot (lapack_info=" +
                                         std::to_string(dg[f].lapack_info) + ")");
            if (diag) diag[f] = dg[f];
        }
        for (size_t v = 0; v < FS; ++v) {
            for (int o = 0; o < n_obs; ++o) {
                field_re[v * n_obs + o] = fh[(v * n_obs + o) * 2];
                field_im[v * n_obs + o] = fh[(v * n_obs + o) * 2 + 1];
            }
            if (want_p)
                for (int i = 0; i < n; ++i) {
                    p_re[v * n + i] = ph[(v * n + i) * 2];
                    p_im[v * n + i] = ph[(v * n + i) * 2 + 1];
                }
            if (impedance) {
                impedance[2 * v] = th[(v * ntag) * 2];
                impedance[2 * v + 1] = th[(v * ntag) * 2 + 1];
            }
            if (pavg)
                for (int t = 0; t < opts->n_tags; ++t) {
                    pavg[(v * opts->n_tags + t) * 2] = th[(v * ntag + 1 + t) * 2];
                    pavg[(v * opts->n_tags + t) * 2 + 1] = th[(v * ntag + 1 + t) * 2 + 1];
                }
        }
    });
}

int32_t hbem_solve_batch_device(hbem_session *s, int32_t F, const float *k_real,
                                const float *k_imag, const float *field_k, const float *q,
                                int32_t n_obs, const float *obs_3xn,
                                const hbem_solve_options *opts, float *field, float *p,
                                double *tag_avg, hbem_case_diag *diag) {
    return guarded([&] {
        if (!s || !opts) fail(HBEM_EINVAL, "hbem_solve_batch_device: null argument");
        if (F <= 0) fail(HBEM_EINVAL, "hbem_solve_batch_device: n_freq must be positive");
        if (!k_real || !field_k || !q || !tag_avg)
            fail(HBEM_EINVAL, "hbem_solve_batch_device: k_real, field_k, q, tag_avg required");
        if (n_obs < 0 || (n_obs > 0 && (!obs_3xn || !field)))
            fail(HBEM_EINVAL, "hbem_solve_batch_device: observation arrays invalid");
        check_options(*opts);
        if (opts->solve_trans != 0)
            fail(HBEM_EINVAL, "hbem_solve_batch_device: solve_trans must be 0");
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        set_device(s);
        std::vector<hbem_case_diag> dg;
        s->chief_pts_d = opts->chief_points;
        solve_core(s, F, k_real, k_imag, field_k, q, opts->robin_beta, n_obs, obs_3xn, *opts,
                   field, p, tag_avg, &dg);
        hip_check(hipStreamSynchronize(s->stream), "batch sync");
        for (int f = 0; f < F; ++f) {
            if (dg[f].lapack_info != 0 && !diag)
                fail(HBEM_ESINGULAR, "LU found an exactly zero pivot");
            if (diag) diag[f] = dg[f];
        }
    });
}

int32_t hbem_adjoint_forward(hbem_session *s, int32_t F, const float *k_real,
                             const float *k_imag, const float *field_k, const float *q,
                             int32_t n_obs, const float *obs_3xn, const hbem_solve_options *opts,
                             float *field, float *p, double *tag_avg, hbem_case_diag *diag) {
    return guarded([&] {
        if (!s || !opts) fail(HBEM_EINVAL, "hbem_adjoint_forward: null argument");
        if (F <= 0) fail(HBEM_EINVAL, "hbem_adjoint_forward: n_freq must be positive");
        if (!k_real || !field_k || !q || !tag_avg || n_obs <= 0 || !obs_3xn || !field)
            fail(HBEM_EINVAL, "hbem_adjoint_forward: k_real, field_k, q, tag_avg, obs, field required");
        check_options(*opts);
        if (opts->solve_trans != 0) fail(HBEM_EINVAL, "hbem_adjoint_forward: solve_trans must be 0");
        if (opts->n_sources > 1)
            fail(HBEM_EINVAL, "hbem_adjoint_forward: one source per frequency");
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        if (s->dg.nc_np > 0)
            fail(HBEM_EINVAL, "hbem_adjoint_forward: near-field quadrature must be off");
        if (s->assembly_mode != HBEM_ASSEMBLY_CORRECTED)
            fail(HBEM_EUNSUPPORTED, "hbem_adjoint_forward: needs the corrected assembly mode");
        set_device(s);
        hipStream_t st = s->stream;
        const int T = s->plan.n_triangles;
        auto keep = [&](DevBuf &d, const float *src, size_t count) {
            d.reserve(sizeof(float) * count);
            hip_check(hipMemcpyAsync(d.ptr, src, sizeof(float) * count, hipMemcpyDeviceToDevice, st),
                      "adjoint input copy");
        };
        keep(s->adj_k, k_real, F);
        keep(s->adj_fk, field_k, F);
        if (k_imag) keep(s->adj_kim, k_imag, F);
        keep(s->adj_q, q, (size_t)2 * F * T);
        keep(s->adj_obs, obs_3xn, (size_t)3 * n_obs);
        if (opts->robin_beta) keep(s->adj_beta, opts->robin_beta, (size_t)2 * F * T);
        hbem_solve_options o = *opts;
        o.n_sources = 1;
        std::vector<hbem_case_diag> dg;
        s->adj_keep = true;
        try {
            solve_core(s, F, s->adj_k.as<float>(), k_imag ? s->adj_kim.as<float>() : nullptr,
                       s->adj_fk.as<float>(), s->adj_q.as<float>(),
                       opts->robin_beta ? s->adj_beta.as<float>() : nullptr, n_obs,
                       s->adj_obs.as<float>(), o, field, p, tag_avg, &dg);
        } catch (...) {
            s->adj_keep = false;
            throw;
        }
        s->adj_keep = false;
        hip_check(hipStreamSynchronize(st), "adjoint forward sync");
        for (int f = 0; f < F; ++f) {
            if (dg[f].lapack_info != 0 && !diag)
                fail(HBEM_ESINGULAR, "LU found an exactly zero pivot");
            if (diag) diag[f] = dg[f];
        }
        s->adj_F = F;
        s->adj_nobs = n_obs;
        s->adj_complex = k_imag != nullptr;
        s->adj_robin = opts->robin_beta != nullptr;
        s->adj_valid = true;
    });
}

int32_t hbem_adjoint_backward(hbem_session *s, const float *grad_field, double *grad_vertices,
                              float *grad_q, int32_t refine_iterations) {
    return guarded([&] {
        if (!s || !grad_field || !grad_vertices)
            fail(HBEM_EINVAL, "hbem_adjoint_backward: null argument");
        std::lock_guard<std::mutex> lock(s->mu);
        if (!s->adj_valid)
            fail(HBEM_EINVAL,
                 "hbem_adjoint_backward: no adjoint forward is resident on this session");
        set_device(s);
        hipStream_t st = s->stream;
        const int B = s->adj_F, n = s->plan.n_dofs, T = s->plan.n_triangles;
        const int nv = s->plan.n_vertices, n_obs = s->adj_nobs;
        const int N = (int)round_up(n, hbem::kLuPad);
        const int64_t lda = N, a_stride = (int64_t)N * N;
        const float *kre = s->adj_k.as<float>();
        const float *kim = s->adj_complex ? s->adj_kim.as<float>() : nullptr;
        const float *qt = s->adj_robin ? s->qt.as<float>() : s->adj_q.as<float>();
        const float *pv = s->x.as<float>();
        s->adj_c.reserve(sizeof(double) * 2 * (size_t)B * N);
        s->adj_ft.reserve(sizeof(double) * 2 * (size_t)B * T);
        s->adj_sig.reserve(sizeof(double) * 2 * (size_t)B * T);
        s->adj_gN.reserve(sizeof(double) * 3 * (size_t)T);
        s->adj_l32.reserve(sizeof(float) * 2 * (size_t)B * N);
        s->adj_e64.reserve(sizeof(double) * 2 * (size_t)B * N);
        s->adj_l64.reserve(sizeof(double) * 2 * (size_t)B * N);
        s->adj_r64.reserve(sizeof(double) * 2 * (size_t)B * N);
        s->norms.reserve(sizeof(double) * 2 * B);
        hbem::launch_fill_zero_f64(grad_vertices, (size_t)3 * nv, st);
        hbem::launch_fill_zero_f64(s->adj_c.as<double>(), (size_t)2 * B * N, st);
        hbem::launch_fill_zero_f64(s->adj_ft.as<double>(), (size_t)2 * B * T, st);
        hbem::launch_fill_zero_f64(s->adj_sig.as<double>(), (size_t)2 * B * T, st);
        hbem::launch_fill_zero_f64(s->adj_gN.as<double>(), (size_t)3 * T, st);
        hbem::launch_field_adjoint(s->dg, B, s->adj_fk.as<float>(), pv, N, qt, n_obs,
                                   s->adj_obs.as<float>(), grad_field, s->adj_c.as<double>(), N,
                                   s->adj_ft.as<double>(), grad_vertices, s->adj_gN.as<double>(), st);
        hbem::launch_adjoint_rhs(s->dg, B, kre, kim, s->adj_robin ? s->adj_beta.as<float>() : nullptr,
                                 s->adj_ft.as<double>(), s->adj_c.as<double>(), N, n, N,
                                 s->adj_l32.as<float>(), s->adj_e64.as<double>(), st);
        // lambda = A^{-H} e with the resident LU, refined against the FP64
        // residual e - A^H lambda of the kept matrix.
        hbem::LuWorkspace w{s->A.as<float>(), lda, a_stride, N, n, B, s->ipiv.as<int32_t>(),
                            s->info.as<int32_t>(), s->gl.as<int32_t>(), s->lu_scr_i.as<int32_t>(),
                            s->lu_scr_f.as<float>()};
        hbem::lu_solve(w, s->adj_l32.as<float>(), N, 2, st);
        hbem::launch_fill_zero_f64(s->adj_l64.as<double>(), (size_t)2 * B * N, st);
        hbem::launch_axpy_f32_to_f64(s->adj_l32.as<float>(), s->adj_l64.as<double>(), N, N, n, B,
                                     nullptr, st);
        for (int it = 0; it < refine_iterations; ++it) {
            hbem::launch_residual_h_f64(s->A_orig.as<float>(), lda, a_stride, n, B,
                                        s->adj_e64.as<double>(), s->adj_l64.as<double>(), N,
                                        s->adj_r64.as<double>(), s->norms.as<double>(), st);
            hbem::launch_fill_zero(s->adj_l32.as<float>(), (size_t)2 * B * N, st);
            hbem::launch_copy_f64_to_f32(s->adj_r64.as<double>(), s->adj_l32.as<float>(), N, N, n,
                                         B, st);
            hbem::lu_solve(w, s->adj_l32.as<float>(), N, 2, st);
            hbem::launch_axpy_f32_to_f64(s->adj_l32.as<float>(), s->adj_l64.as<double>(), N, N, n,
                                         B, nullptr, st);
        }
        hbem::launch_fill_zero(s->adj_l32.as<float>(), (size_t)2 * B * N, st);
        hbem::launch_copy_f64_to_f32(s->adj_l64.as<double>(), s->adj_l32.as<float>(), N, N, n, B,
                                     st);
        hbem::launch_assembly_adjoint(s->dg, B, kre, kim, s->adj_l32.as<float>(), pv, N, qt,
                                      grad_vertices, s->adj_gN.as<double>(),
                                      s->adj_sig.as<double>(), st);
        hbem::launch_adjoint_finish(s->dg, B, s->adj_l32.as<float>(), pv, N,
                                    s->adj_gN.as<double>(), grad_vertices, s->adj_ft.as<double>(),
                                    s->adj_sig.as<double>(), grad_q, st);
        hip_check(hipGetLastError(), "adjoint launch");
        hip_check(hipStreamSynchronize(st), "adjoint sync");
    });
}

int32_t hbem_evaluate_field(hbem_session *s, int32_t F, const float *field_k, const float *p_re,
                            const float *p_im, const float *q_re, const float *q_im, int32_t n_obs,
                            const float *obs_3xn, float *out_re, float *out_im) {
    return guarded([&] {
        if (!s) fail(HBEM_EINVAL, "hbem_evaluate_field: null session");
        if (F <= 0 || n_obs <= 0) fail(HBEM_EINVAL, "hbem_evaluate_field: empty batch");
        if (!field_k || !p_re || !p_im || !q_re || !q_im || !obs_3xn || !out_re || !out_im)
            fail(HBEM_EINVAL, "hbem_evaluate_field: null array");
        std::lock_guard<std::mutex> lock(s->mu);
        s->adj_valid = false;
        set_device(s);
        hipStream_t st = s->stream;
        const int n = s->plan.n_dofs, T = s->plan.n_triangles;
        std::vector<float> fk(field_k, field_k + F);
        upload(s->fk, fk, st);
        std::vector<float> p = interleave(p_re, p_im, (size_t)F * n);
        upload(s->x, p, st);
        std::vector<float> q = interleave(q_re, q_im, (size_t)F * T);
        upload(s->q, q, st);
        std::vector<float> obs(obs_3xn, obs_3xn + 3 * (size_t)n_obs);
        upload(s->obs, obs, st);
        s->field.reserve(sizeof(float) * 2 * (size_t)F * n_obs);
        hbem::launch_fill_zero(s->field.as<float>(), (size_t)2 * F * n_obs, st);
        hbem::launch_field(s->dg, F, s->fk.as<float>(), 1, s->x.as<float>(), n, s->q.as<float>(),
                           n_obs, s->obs.as<float>(), s->field.as<float>(), st);
        hip_check(hipGetLastError(), "field launch");
        std::vector<float> fh((size_t)2 * F * n_obs);
        hip_check(hipMemcpyAsync(fh.data(), s->field.ptr, sizeof(float) * fh.size(),
                                 hipMemcpyDeviceToHost, st),
                  "field D2H");
        hip_check(hipStreamSynchronize(st), "field sync");
        for (size_t i = 0; i < (size_t)F * n_obs; ++i) {
            out_re[i] = fh[2 * i];
            out_im[i] = fh[2 * i + 1];
        }
    });
}

int32_t hbem_dense_solve(int32_t n_sys, int32_t n, const float *A_re, const float *A_im,
                         const float *b_re, const float *b_im, const hbem_solve_options *opts,
                         float *x_re, float *x_im, hbem_case_diag *diag) {
    return guarded([&] {
        if (n_sys <= 0 || n <= 0) fail(HBEM_EINVAL, "hbem_dense_solve: empty problem");
        if (!A_re || !A_im || !b_re || !b_im || !x_re || !x_im || !opts)
            fail(HBEM_EINVAL, "hbem_dense_solve: null array");
        check_options(*opts);
        if (opts->n_chief > 0)
            fail(HBEM_EINVAL, "hbem_dense_solve: CHIEF rows need an assembled session");
        if (opts->n_sources > 1)
            fail(HBEM_EINVAL, "hbem_dense_solve: one right-hand side per system");
        if ((int64_t)n > hbem::kLuMaxOrder)
            fail(HBEM_EINVAL, "hbem_dense_solve: order above the supported maximum (32768)");
        const int dev = gfx950_device();
        if (dev < 0) fail(HBEM_ENODEV, "no gfx950 (MI355X) device is visible");
        // A transient session without geometry carries the workspaces.
        hbem_session s;
        s.device = dev;
        set_device(&s);
        hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
        const int N = (int)round_up(n, hbem::kLuPad);
        const int64_t lda = N, a_stride = (int64_t)N * N;
        const int B = n_sys;
        s.A.reserve(sizeof(float) * 2 * (size_t)B * a_stride);
        s.rowsplit.reserve(sizeof(float) * 2 * (size_t)B * n * n);
        float *re_d = s.rowsplit.as<float>();
        float *im_d = re_d + (size_t)B * n * n;
        hip_check(hipMemcpyAsync(re_d, A_re, sizeof(float) * (size_t)B * n * n,
                                 hipMemcpyHostToDevice, s.stream),
                  "A_re H2D");
        hip_check(hipMemcpyAsync(im_d, A_im, sizeof(float) * (size_t)B * n * n,
                                 hipMemcpyHostToDevice, s.stream),
                  "A_im H2D");
        hbem::launch_fill_zero(s.A.as<float>(), (size_t)2 * B * a_stride, s.stream);
        hbem::launch_rowmajor_split_to_colmajor(re_d, im_d, n, B, s.A.as<float>(), lda, a_stride,
                                                s.stream);
        std::vector<float> rh((size_t)2 * B * N, 0.0f);
        for (int b = 0; b < B; ++b)
            for (int i = 0; i < n; ++i) {
                rh[((size_t)b * N + i) * 2] = b_re[(size_t)b * n + i];
                rh[((size_t)b * N + i) * 2 + 1] = b_im[(size_t)b * n + i];
            }
        upload(s.rhs, rh, s.stream);
        std::vector<hbem_case_diag> dg;
        double solve_s = 0.0;
        solve_resident(&s, B, 1, n, N, *opts, &dg, &solve_s);
        std::vector<float> xh((size_t)2 * B * N);
        hip_check(hipMemcpyAsync(xh.data(), s.x.ptr, sizeof(float) * xh.size(),
                                 hipMemcpyDeviceToHost, s.stream),
                  "x D2H");
        hip_check(hipStreamSynchronize(s.stream), "dense solve sync");
        for (int b = 0; b < B; ++b) {
            if (dg[b].lapack_info != 0 && !diag)
                fail(HBEM_ESINGULAR, "LU found an exactly zero pivot");
            for (int i = 0; i < n; ++i) {
                x_re[(size_t)b * n + i] = xh[((size_t)b * N + i) * 2];
                x_im[(size_t)b * n + i] = xh[((size_t)b * N + i) * 2 + 1];
            }
            if (diag) {
                diag[b] = dg[b];
                diag[b].assembly_s = 0.0;
                diag[b].field_s = 0.0;
                diag[b].dense_solve_s = solve_s / B;
            }
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Session groups: several meshes' sweeps solved together (the batched uneven
// sweep).  One regular and one singular assembly launch cover every member;
// the members' systems are padded into size classes of a common order so
// that each class factors and solves as one batch; each class runs on its own
// stream, forked from and joined to the group stream; the whole sequence is
// captured once per shape and buffer set and replayed as one hipGraph.
// ---------------------------------------------------------------------------

struct hbem_group {
    std::vector<hbem_session *> members;
    int device = -1;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> side;   // graph capture: one per size class
    std::vector<hipStream_t> lanes;  // per-class assembly: streams with hardware queues of their own
    hipStream_t asm_lane = nullptr;  // per-class assembly: the assembly launches (own queue)
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;  // per class: solves joined, assembly done (split)
    struct SizeClass {
        int N = 0;
        int B = 0;
        std::vector<int> members, offset;
        DevBuf A, rhs, x, ipiv, info, gl, scr_i, scr_f, nsys;
    };
    std::vector<SizeClass> classes;
    int granularity = 512;   // padded-order step of the size classes (autotuned)
    int classes_gran = 0;    // the step the current classes were built with
    bool split = false;      // one assembly launch per size class (autotuned)
    bool tuned = false;
    DevBuf jobs_d, tags_d, stamps;  // stamps: [0] sweep start, [1 + m] member m done
    double clock_hz = 0.0;
    hipGraphExec_t graph = nullptr;
    std::vector<uintptr_t> graph_key;

    ~hbem_group() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (SizeClass &c : classes)
            for (DevBuf *b : {&c.A, &c.rhs, &c.x, &c.ipiv, &c.info, &c.gl, &c.scr_i, &c.scr_f, &c.nsys})
                b->release();
        jobs_d.release();
        tags_d.release();
        stamps.release();
        for (hipEvent_t e : ev_join) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        for (hipStream_t s : lanes) (void)hipStreamDestroy(s);
        if (asm_lane) (void)hipStreamDestroy(asm_lane);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int kGroupGranularities[4] = {64, 256, 512, 1024};
constexpr int kGroupCandidates = 8;  // granularities x {one assembly launch, one per class}
// Streams the per-class sweep spreads the size classes' solves over.  They and
// the stream of the assembly launches are created with a CU mask (all CUs),
// which gives each a hardware queue of its own: streams created without one
// share the process's few hardware queues, where a solve queued between two
// assembly launches delays the second.  (Reserving CUs for the solves, or
// 2, 4 or 6 solve streams, measured slower on the 16-mesh sweep.)
constexpr int kGroupLanes = 3;

// CU mask selecting all `total` CUs.
std::vector<uint32_t> full_cu_mask(int total) {
    std::vector<uint32_t> mask((size_t)(total + 31) / 32, 0u);
    for (int i = 0; i < total; ++i) mask[(size_t)i / 32] |= 1u << (i % 32);
    return mask;
}

void check_group_case(const hbem_session *s, const hbem_group_case &c) {
    if (c.n_freq <= 0) fail(HBEM_EINVAL, "hbem_group: n_freq must be positive");
    if (!c.k_real || !c.field_k || !c.q || !c.tag_avg)
        fail(HBEM_EINVAL, "hbem_group: k_real, field_k, q and tag_avg are required");
    if (c.n_obs < 0 || (c.n_obs > 0 && (!c.obs_3xn || !c.field)))
        fail(HBEM_EINVAL, "hbem_group: observation arrays invalid");
    (void)s;
}

// Size classes of padded order round_up(N_m, granularity), largest first.
void build_classes(hbem_group *g, const hbem_group_case *cases, int gran) {
    std::vector<std::pair<int, int>> order;  // (class order, member)
    for (size_t m = 0; m < g->members.size(); ++m) {
        const int N = (int)round_up(g->members[m]->plan.n_dofs, hbem::kLuPad);
        order.push_back({(int)round_up(N, gran), (int)m});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const auto &a, const auto &b) { return a.first > b.first; });
    std::vector<hbem_group::SizeClass> keep = std::move(g->classes);
    g->classes.clear();
    for (const auto &om : order) {
        if (g->classes.empty() || g->classes.back().N != om.first) {
            g->classes.emplace_back();
            g->classes.back().N = om.first;
        }
        hbem_group::SizeClass &c = g->classes.back();
        c.members.push_back(om.second);
        c.offset.push_back(c.B);
        c.B += cases[om.second].n_freq;
    }
    // reuse the previous classes' buffers where they exist
    for (size_t i = 0; i < g->classes.size() && i < keep.size(); ++i) {
        hbem_group::SizeClass &c = g->classes[i];
        c.A = keep[i].A; c.rhs = keep[i].rhs; c.x = keep[i].x; c.ipiv = keep[i].ipiv;
        c.info = keep[i].info; c.gl = keep[i].gl; c.scr_i = keep[i].scr_i;
        c.scr_f = keep[i].scr_f; c.nsys = keep[i].nsys;
        keep[i].A = keep[i].rhs = keep[i].x = keep[i].ipiv = keep[i].info = keep[i].gl =
            keep[i].scr_i = keep[i].scr_f = keep[i].nsys = DevBuf{};
    }
    for (hbem_group::SizeClass &c : keep)
        for (DevBuf *b : {&c.A, &c.rhs, &c.x, &c.ipiv, &c.info, &c.gl, &c.scr_i, &c.scr_f, &c.nsys})
            b->release();
    for (hbem_group::SizeClass &c : g->classes) {
        const size_t N = (size_t)c.N, B = (size_t)c.B;
        c.A.reserve(sizeof(float) * 2 * B * N * N);
        c.rhs.reserve(sizeof(float) * 2 * B * N);
        c.x.reserve(sizeof(float) * 2 * B * N);
        c.ipiv.reserve(sizeof(int32_t) * B * N);
        c.info.reserve(sizeof(int32_t) * B);
        c.gl.reserve(sizeof(int32_t) * B * hbem::lu_gather_ints(c.N));
        c.scr_i.reserve(sizeof(int32_t) * B * hbem::kLuScratchI);
        c.scr_f.reserve(sizeof(float) * B * hbem::kLuScratchF);
        std::vector<int32_t> nsys;
        for (size_t i = 0; i < c.members.size(); ++i)
            for (int f = 0; f < cases[c.members[i]].n_freq; ++f)
                nsys.push_back(g->members[c.members[i]]->plan.n_dofs);
        upload(c.nsys, nsys, g->stream);
    }
}

// Enqueues the whole group sweep on g->stream, forking the size classes'
// solves onto g->side (one stream per class, for graph capture) or, when the
// assembly is split per class, onto g->lanes (launched directly: a replayed
// graph runs the per-class sweep's branches one after another on one
// hardware queue).  Kernels and device copies only.
// jobs: per size class, the class's members (heaviest first), classes in
// order; class_job[c] = index of the class's first job (class_job[C] = jobs).
void enqueue_group(hbem_group *g, const hbem_group_case *cases, const std::vector<hbem::AssemblyJob> &jobs,
                   const std::vector<int> &class_job, int ntag) {
    hipStream_t st = g->stream;
    long long *stamps = g->stamps.as<long long>();
    hbem::launch_timestamp(stamps, st);
    for (size_t m = 0; m < g->members.size(); ++m)
        if (cases[m].n_obs > 0)
            hbem::launch_fill_zero(cases[m].field, (size_t)2 * cases[m].n_freq * cases[m].n_obs, st);
    for (hbem_group::SizeClass &c : g->classes) {
        hbem::launch_fill_zero(c.A.as<float>(), (size_t)2 * c.B * c.N * c.N, st);
        hbem::launch_fill_zero(c.rhs.as<float>(), (size_t)2 * c.B * c.N, st);
    }
    hipStream_t as = st;  // the assembly's stream
    if (!g->split) {
        // one launch over every member's tile pairs, one over every member's
        // singular classes; the size classes' solves fork after it
        hbem::launch_assemble_multi(g->jobs_d.as<hbem::AssemblyJob>(), jobs, false, st);
        hip_check(hipEventRecord(g->ev_fork, st), "event record");
    } else {
        hip_check(hipEventRecord(g->ev_fork, st), "event record");
        as = g->asm_lane;
        hip_check(hipStreamWaitEvent(as, g->ev_fork, 0), "fork");
    }
    for (size_t ci = 0; ci < g->classes.size(); ++ci) {
        hbem_group::SizeClass &c = g->classes[ci];
        hipStream_t ss = g->split ? g->lanes[ci % g->lanes.size()] : g->side[ci];
        if (g->split) {
            // the class's members in one launch each; its solves fork while
            // the next class assembles
            const std::vector<hbem::AssemblyJob> cj(jobs.begin() + class_job[ci],
                                                    jobs.begin() + class_job[ci + 1]);
            hbem::launch_assemble_multi(g->jobs_d.as<hbem::AssemblyJob>() + class_job[ci], cj, false,
                                        as);
            hip_check(hipEventRecord(g->ev_join[ci + g->classes.size()], as), "event record");
            hip_check(hipStreamWaitEvent(ss, g->ev_join[ci + g->classes.size()], 0), "fork");
        } else {
            hip_check(hipStreamWaitEvent(ss, g->ev_fork, 0), "fork");
        }
        const int64_t N = c.N, a_stride = N * N;
        hbem::launch_pad_identity_var(c.A.as<float>(), N, a_stride, c.nsys.as<int32_t>(), c.N, c.B,
                                      ss);
        hbem::LuWorkspace w{c.A.as<float>(), N, a_stride, c.N, c.N, c.B, c.ipiv.as<int32_t>(),
                            c.info.as<int32_t>(), c.gl.as<int32_t>(), c.scr_i.as<int32_t>(),
                            c.scr_f.as<float>()};
        hbem::lu_factor(w, 0, nullptr, ss);
        hip_check(hipMemcpyAsync(c.x.ptr, c.rhs.ptr, sizeof(float) * 2 * (size_t)c.B * c.N,
                                 hipMemcpyDeviceToDevice, ss),
                  "rhs copy");
        hbem::lu_solve(w, c.x.as<float>(), N, 0, ss);
        for (size_t i = 0; i < c.members.size(); ++i) {
            const int m = c.members[i];
            const hbem_session *s = g->members[m];
            const hbem_group_case &cs = cases[m];
            const float *xm = c.x.as<float>() + (size_t)c.offset[i] * N * 2;
            if (cs.n_obs > 0)
                hbem::launch_field(s->dg, cs.n_freq, cs.field_k, 1, xm, N, cs.q, cs.n_obs,
                                   cs.obs_3xn, cs.field, ss);
            hbem::launch_tag_average(s->dg, cs.n_freq, xm, N, ntag, g->tags_d.as<int32_t>(),
                                     cs.tag_avg, ss);
            if (cs.p)
                hip_check(hipMemcpy2DAsync(cs.p, sizeof(float) * 2 * s->plan.n_dofs, xm,
                                           sizeof(float) * 2 * N,
                                           sizeof(float) * 2 * s->plan.n_dofs, cs.n_freq,
                                           hipMemcpyDeviceToDevice, ss),
                          "p copy");
            hbem::launch_timestamp(stamps + 1 + m, ss);
        }
        hip_check(hipEventRecord(g->ev_join[ci], ss), "event record");
        hip_check(hipStreamWaitEvent(st, g->ev_join[ci], 0), "join");
    }
    if (as != st) {
        // the last class's assembly event, already recorded on as
        hip_check(hipStreamWaitEvent(st, g->ev_join[2 * g->classes.size() - 1], 0), "join");
    }
    hip_check(hipGetLastError(), "group launch");
}

// One group sweep at the given granularity.  Fills diag (member order,
// frequency order within a member) and member_seconds (completion time of
// each member's results after the sweep started).
void group_solve(hbem_group *g, const hbem_group_case *cases, const hbem_solve_options &opts,
                 int gran, hbem_case_diag *diag, double *member_seconds, double *sweep_seconds) {
    const size_t M = g->members.size();
    hipStream_t st = g->stream;
    bool rebuild = gran != g->classes_gran || g->classes.empty();
    for (const hbem_group::SizeClass &c : g->classes) {
        int B = 0;
        for (int m : c.members) B += cases[m].n_freq;
        rebuild |= B != c.B;  // member batch sizes changed
    }
    if (rebuild) {
        build_classes(g, cases, gran);
        g->classes_gran = gran;
    }
    while (!g->split && g->side.size() < g->classes.size()) {
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        g->side.push_back(s);
    }
    if (g->split && g->lanes.empty()) {
        int ncu = 0;
        hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->device),
                  "CU count");
        const std::vector<uint32_t> mask = full_cu_mask(ncu);
        for (int i = 0; i < kGroupLanes; ++i) {
            hipStream_t s = nullptr;
            hip_check(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()),
                      "hipExtStreamCreateWithCUMask");
            g->lanes.push_back(s);
        }
        hip_check(hipExtStreamCreateWithCUMask(&g->asm_lane, (uint32_t)mask.size(), mask.data()),
                  "hipExtStreamCreateWithCUMask");
    }
    while (g->ev_join.size() < 2 * g->classes.size()) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        g->ev_join.push_back(e);
    }
    g->stamps.reserve(sizeof(long long) * (1 + M));
    if (g->clock_hz == 0.0) {
        int khz = 0;
        hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device),
                  "wall clock rate");
        g->clock_hz = 1e3 * (double)khz;
    }
    // assembly jobs: by size class, within a class the members with the most
    // blocks first; block offsets restart at every class when the assembly
    // is split per class
    std::vector<hbem::AssemblyJob> jobs;
    std::vector<int> class_job;
    int reg = 0, sing = 0;
    for (size_t ci = 0; ci < g->classes.size(); ++ci) {
        const hbem_group::SizeClass &c = g->classes[ci];
        class_job.push_back((int)jobs.size());
        if (g->split) reg = sing = 0;
        std::vector<int> by_work(c.members.size());
        for (size_t i = 0; i < by_work.size(); ++i) by_work[i] = (int)i;
        std::stable_sort(by_work.begin(), by_work.end(), [&](int a, int b) {
            const int ma = c.members[a], mb = c.members[b];
            return (int64_t)g->members[ma]->dg.n_tile_pairs * cases[ma].n_freq >
                   (int64_t)g->members[mb]->dg.n_tile_pairs * cases[mb].n_freq;
        });
        for (int i : by_work) {
        const int m = c.members[i], oi = c.offset[i];
        const hbem_session *s = g->members[m];
        hbem::AssemblyJob j{};
        j.g = s->dg;
        j.k_real = cases[m].k_real;
        j.k_imag = nullptr;
        j.q = cases[m].q;
        j.beta = nullptr;
        j.n_src = 1;
        j.n_batch = cases[m].n_freq;
        j.lda = c.N;
        j.a_stride = (int64_t)c.N * c.N;
        j.A = c.A.as<float>() + (size_t)oi * j.a_stride * 2;
        j.rhs_stride = c.N;
        j.rhs = c.rhs.as<float>() + (size_t)oi * c.N * 2;
        j.block_begin_reg = reg;
        j.block_begin_sing = sing;
        reg += hbem::regular_blocks(s->dg, j.n_batch);
        sing += hbem::singular_blocks(s->dg, j.n_batch);
        jobs.push_back(j);
        }
    }
    class_job.push_back((int)jobs.size());
    upload(g->jobs_d, jobs, st);
    std::vector<int32_t> tag_list;
    tag_list.push_back(opts.impedance_tag);
    for (int i = 0; i < opts.n_tags; ++i) tag_list.push_back(opts.tags[i]);
    upload(g->tags_d, tag_list, st);
    const int ntag = (int)tag_list.size();
    std::vector<uintptr_t> key = {(uintptr_t)gran, (uintptr_t)g->split, (uintptr_t)g->jobs_d.ptr,
                                  (uintptr_t)g->tags_d.ptr, (uintptr_t)g->stamps.ptr,
                                  (uintptr_t)ntag};
    for (size_t m = 0; m < M; ++m) {
        const hbem_group_case &c = cases[m];
        for (uintptr_t v : {(uintptr_t)c.n_freq, (uintptr_t)c.k_real, (uintptr_t)c.field_k,
                            (uintptr_t)c.q, (uintptr_t)c.n_obs, (uintptr_t)c.obs_3xn,
                            (uintptr_t)c.field, (uintptr_t)c.p, (uintptr_t)c.tag_avg})
            key.push_back(v);
    }
    for (const hbem_group::SizeClass &c : g->classes)
        for (const DevBuf *b : {&c.A, &c.rhs, &c.x, &c.ipiv, &c.info, &c.gl, &c.nsys})
            key.push_back((uintptr_t)b->ptr);
    if (!g->split && (!g->graph || g->graph_key != key)) {
        if (g->graph) {
            (void)hipGraphExecDestroy(g->graph);
            g->graph = nullptr;
        }
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture begin");
        hipGraph_t graph = nullptr;
        try {
            enqueue_group(g, cases, jobs, class_job, ntag);
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hip_check(hipStreamEndCapture(st, &graph), "capture end");
        const hipError_t e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(e, "graph instantiate");
        g->graph_key = key;
    }
    Timer tw;
    tw.start(st);
    if (g->split)
        enqueue_group(g, cases, jobs, class_job, ntag);
    else
        hip_check(hipGraphLaunch(g->graph, st), "graph launch");
    tw.stop(st);
    const double sweep_s = tw.seconds();
    if (sweep_seconds) *sweep_seconds = sweep_s;
    // per-class pivot info; member completion times from the device stamps
    std::vector<long long> stamp(1 + M);
    hip_check(hipMemcpy(stamp.data(), g->stamps.ptr, sizeof(long long) * (1 + M),
                        hipMemcpyDeviceToHost),
              "stamps D2H");
    size_t d0 = 0;
    std::vector<size_t> first(M);
    for (size_t m = 0; m < M; ++m) {
        first[m] = d0;
        d0 += cases[m].n_freq;
    }
    for (size_t ci = 0; ci < g->classes.size(); ++ci) {
        const hbem_group::SizeClass &c = g->classes[ci];
        std::vector<int32_t> info(c.B);
        hip_check(hipMemcpy(info.data(), c.info.ptr, sizeof(int32_t) * c.B, hipMemcpyDeviceToHost),
                  "info D2H");
        for (size_t i = 0; i < c.members.size(); ++i) {
            const int m = c.members[i];
            const double done = (double)(stamp[1 + m] - stamp[0]) / g->clock_hz;
            if (member_seconds) member_seconds[m] = done;
            const double per = done / cases[m].n_freq;
            for (int f = 0; f < cases[m].n_freq; ++f) {
                hbem_case_diag &d = diag[first[m] + f];
                d = hbem_case_diag{};
                d.lapack_info = info[c.offset[i] + f];
                d.dense_solve_rcond = -1.0;
                d.refine_residual_rel = -1.0;
                d.chief_residual_rel = -1.0;
                d.dense_solve_s = per;
                d.batch_index = (int)ci;
                d.batch_size = c.B;
            }
        }
    }
}

// Locks every member session (address order) for the duration of a call.
struct GroupLock {
    std::vector<std::unique_lock<std::mutex>> locks;
    explicit GroupLock(hbem_group *g) {
        std::vector<hbem_session *> m = g->members;
        std::sort(m.begin(), m.end());
        for (hbem_session *s : m) locks.emplace_back(s->mu);
    }
};

void check_group_options(const hbem_solve_options &o) {
    check_options(o);
    if (o.solve_mode != HBEM_SOLVE_F32 || o.estimate_rcond || o.robin_beta || o.n_chief > 0 ||
        o.n_sources > 1 || o.precise_geometry || o.solve_trans != 0)
        fail(HBEM_EUNSUPPORTED,
             "hbem_group: FP32 LU without rcond, Robin walls, CHIEF rows, extra sources or "
             "precise geometry");
}

}  // namespace

extern "C" {

int32_t hbem_group_create(int32_t n_sessions, hbem_session *const *sessions, hbem_group **out) {
    return guarded([&] {
        if (!out || !sessions || n_sessions <= 0 || n_sessions > 1024)
            fail(HBEM_EINVAL, "hbem_group_create: invalid arguments");
        *out = nullptr;
        auto g = std::make_unique<hbem_group>();
        for (int i = 0; i < n_sessions; ++i) {
            hbem_session *s = sessions[i];
            if (!s) fail(HBEM_EINVAL, "hbem_group_create: null session");
            for (hbem_session *t : g->members)
                if (t == s) fail(HBEM_EINVAL, "hbem_group_create: a session appears twice");
            if (i > 0 && s->device != sessions[0]->device)
                fail(HBEM_EINVAL, "hbem_group_create: sessions on different devices");
            if (s->dg.n_masks != sessions[0]->dg.n_masks)
                fail(HBEM_EUNSUPPORTED, "hbem_group_create: members need the same symmetry image count");
            g->members.push_back(s);
        }
        g->device = sessions[0]->device;
        hip_check(hipSetDevice(g->device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming), "hipEventCreate");
        *out = g.release();
    });
}

int32_t hbem_group_destroy(hbem_group *group) {
    return guarded([&] {
        if (!group) return;
        (void)hipSetDevice(group->device);
        delete group;
    });
}

int32_t hbem_group_solve_device(hbem_group *group, const hbem_group_case *cases,
                                const hbem_solve_options *opts, hbem_case_diag *diag,
                                double *member_seconds, double *sweep_seconds) {
    return guarded([&] {
        if (!group || !cases || !opts || !diag) fail(HBEM_EINVAL, "hbem_group_solve_device: null argument");
        check_group_options(*opts);
        for (size_t m = 0; m < group->members.size(); ++m) {
            check_group_case(group->members[m], cases[m]);
            if (group->members[m]->dg.nc_np > 0)
                fail(HBEM_EUNSUPPORTED, "hbem_group_solve_device: near-field quadrature must be off");
            if (group->members[m]->assembly_mode != HBEM_ASSEMBLY_CORRECTED)
                fail(HBEM_EUNSUPPORTED, "hbem_group_solve_device: needs the corrected assembly mode");
        }
        GroupLock lock(group);
        for (hbem_session *s : group->members) s->adj_valid = false;
        hip_check(hipSetDevice(group->device), "hipSetDevice");
        group_solve(group, cases, *opts, group->granularity, diag, member_seconds, sweep_seconds);
    });
}

int32_t hbem_group_autotune(hbem_group *group, const hbem_group_case *cases,
                            const hbem_solve_options *opts, int32_t *granularity,
                            double *seconds) {
    return guarded([&] {
        if (!group || !cases || !opts) fail(HBEM_EINVAL, "hbem_group_autotune: null argument");
        check_group_options(*opts);
        for (size_t m = 0; m < group->members.size(); ++m) check_group_case(group->members[m], cases[m]);
        GroupLock lock(group);
        hip_check(hipSetDevice(group->device), "hipSetDevice");
        size_t total = 0;
        for (size_t m = 0; m < group->members.size(); ++m) total += cases[m].n_freq;
        std::vector<hbem_case_diag> diag(total);
        int best = 0;
        double best_s = 1e300;
        for (int i = 0; i < kGroupCandidates; ++i) {
            const int gran = kGroupGranularities[i % 4];
            group->split = i >= 4;
            double t = 0.0, t2 = 0.0;
            // one sweep captures and warms the graph, the next is timed
            group_solve(group, cases, *opts, gran, diag.data(), nullptr, &t);
            group_solve(group, cases, *opts, gran, diag.data(), nullptr, &t2);
            if (seconds) seconds[i] = t2;
            if (t2 < best_s) {
                best_s = t2;
                best = i;
            }
        }
        group->granularity = kGroupGranularities[best % 4];
        group->split = best >= 4;
        group->tuned = true;
        if (granularity) *granularity = group->granularity * (group->split ? -1 : 1);
    });
}

}  // extern "C"
