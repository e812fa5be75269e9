// hbem_dense_solve_noalloc: the dense solve of hbem_dense_solve on device
// arrays with no device allocation, for callers that cannot allocate (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
// Its working storage (the column-major copy of the matrices, pivots,
// composed exchange lists, panel scratch, the solution vectors) is one
// file-scope device array of fixed size, `scratch`, which every GPU the
// library runs on holds until the process ends.  The kernels are the ones
// hbem_dense_solve runs.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/hbem.h"
#include "hbem_internal.h"

namespace {

using hbem::kLuGather;
using hbem::kLuPad;
using hbem::kLuScratchF;
using hbem::kLuScratchI;

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Limits: n_sys * N^2 <= 2^22 complex entries (N = n rounded up to 64), so
// n_sys * N <= 2^16; at most 64 systems and order 2048.
constexpr size_t kMaxEntries = size_t(1) << 22;
constexpr size_t kMaxVecEntries = kMaxEntries / kLuPad;
constexpr int kMaxSystems = 64;
constexpr int kMaxOrder = 2048;

// Byte offsets of the pieces of `scratch`.
constexpr size_t kOffA = 0;
constexpr size_t kOffX = kOffA + align256(kMaxEntries * 8);
constexpr size_t kOffPiv = kOffX + align256(kMaxVecEntries * 8);
constexpr size_t kOffInfo = kOffPiv + align256(kMaxVecEntries * 4);
constexpr size_t kOffGl = kOffInfo + align256(kMaxSystems * 4);
constexpr size_t kOffScrI = kOffGl + align256(kMaxVecEntries / 32 * kLuGather * 4);
constexpr size_t kOffScrF = kOffScrI + align256(kMaxSystems * kLuScratchI * 4);
constexpr size_t kScratchBytes = kOffScrF + align256(kMaxSystems * kLuScratchF * 4);
static_assert(kScratchBytes == HBEM_NOALLOC_SCRATCH_BYTES, "header states the scratch size");

// x [b][N] interleaved complex from split (re, im) [b][n] vectors, zero in
// the padding rows.
__global__ __launch_bounds__(256) void k_na_load_rhs(const float *re, const float *im, int n,
                                                     int N, float *x) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const bool in = i < n;
    const size_t s = (size_t)b * n + i;
    *reinterpret_cast<float2 *>(x + ((size_t)b * N + i) * 2) =
        make_float2(in ? re[s] : 0.0f, in ? im[s] : 0.0f);
}

// split (re, im) [b][n] from x [b][N].
__global__ __launch_bounds__(256) void k_na_store_x(const float *x, int n, int N, float *re,
                                                    float *im) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 v = *reinterpret_cast<const float2 *>(x + ((size_t)b * N + i) * 2);
    re[(size_t)b * n + i] = v.x;
    im[(size_t)b * n + i] = v.y;
}

}  // namespace

// The working storage.  In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the per-device address that
// hipGetSymbolAddress returns (scratch_address below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};   // per device, resolved once with that device current
bool g_scratch_arch[64] = {};   // device checked to be gfx950

unsigned char *scratch_address(int dev) {
    if (dev < 0 || dev >= 64) hbem::fail_with(HBEM_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_arch[dev]) {
        hipDeviceProp_t prop;
        hbem::check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
            hbem::fail_with(HBEM_ENODEV, "the current device is not a gfx950 (MI355X)");
        g_scratch_arch[dev] = true;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hbem::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!p) hbem::fail_with(HBEM_EHIP, "hipGetSymbolAddress returned no address");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

}  // namespace

extern "C" int32_t hbem_dense_solve_noalloc(int32_t n_sys, int32_t n, const float *A_re,
                                            const float *A_im, const float *b_re,
                                            const float *b_im, const hbem_solve_options *opts,
                                            float *x_re, float *x_im, int32_t *info,
                                            void *stream) {
    return hbem::run_guarded([&] {
        if (n_sys <= 0 || n <= 0) hbem::fail_with(HBEM_EINVAL, "hbem_dense_solve_noalloc: empty problem");
        if (!A_re || !A_im || !b_re || !b_im || !x_re || !x_im)
            hbem::fail_with(HBEM_EINVAL, "hbem_dense_solve_noalloc: null array");
        const int64_t N = (n + kLuPad - 1) / kLuPad * kLuPad;
        if (n_sys > kMaxSystems || n > kMaxOrder || (int64_t)n_sys * N * N > (int64_t)kMaxEntries)
            hbem::fail_with(HBEM_EINVAL,
                            "hbem_dense_solve_noalloc: above the scratch limits (n_sys <= 64, "
                            "n <= 2048, n_sys * N^2 <= 2^22 with N = n rounded up to 64); "
                            "use hbem_dense_solve");
        int trans = 0;
        if (opts) {
            if (opts->solve_mode != HBEM_SOLVE_F32 || opts->estimate_rcond != 0 ||
                opts->n_chief != 0 || opts->n_sources > 1)
                hbem::fail_with(HBEM_EINVAL,
                                "hbem_dense_solve_noalloc: FP32 LU only (solve_mode HBEM_SOLVE_F32, "
                                "no rcond, no CHIEF rows, one right-hand side)");
            if (opts->solve_trans != 0 && opts->solve_trans != 2)
                hbem::fail_with(HBEM_EINVAL, "hbem_dense_solve_noalloc: solve_trans must be 0 or 2");
            trans = opts->solve_trans;
        }
        int dev = -1;
        hbem::check_hip(hipGetDevice(&dev), "hipGetDevice");
        unsigned char *base = scratch_address(dev);
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int B = n_sys;
        const int64_t a_stride = N * N;
        float *A = reinterpret_cast<float *>(base + kOffA);
        float *x = reinterpret_cast<float *>(base + kOffX);
        hbem::LuWorkspace w{A,
                            N,
                            a_stride,
                            (int)N,
                            n,
                            B,
                            reinterpret_cast<int32_t *>(base + kOffPiv),
                            reinterpret_cast<int32_t *>(base + kOffInfo),
                            reinterpret_cast<int32_t *>(base + kOffGl),
                            reinterpret_cast<int32_t *>(base + kOffScrI),
                            reinterpret_cast<float *>(base + kOffScrF)};
        // the previous call's contents are overwritten: every piece read
        // below is written first in this call
        hbem::launch_fill_zero(A, (size_t)2 * B * a_stride, st);
        hbem::launch_rowmajor_split_to_colmajor(A_re, A_im, n, B, A, N, a_stride, st);
        hbem::launch_pad_identity(A, N, a_stride, n, (int)N, B, st);
        hbem::lu_factor(w, 0, nullptr, st);
        hipLaunchKernelGGL(k_na_load_rhs, dim3((unsigned)((N + 255) / 256), B), dim3(256), 0, st,
                           b_re, b_im, n, (int)N, x);
        hbem::lu_solve(w, x, N, trans, st);
        hipLaunchKernelGGL(k_na_store_x, dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, st, x,
                           n, (int)N, x_re, x_im);
        if (info)
            hbem::check_hip(hipMemcpyAsync(info, w.info, sizeof(int32_t) * B,
                                           hipMemcpyDeviceToDevice, st),
                            "info copy");
        hbem::check_hip(hipGetLastError(), "hbem_dense_solve_noalloc launch");
    });
}
