// C ABI of include/mpmc_hnu_amd.h: mixture residency, batched flash and batched
// phase evaluation. Argument checks mirror the reference's input contracts; numerical
// outcomes are per-cell statuses written by the kernels.
#include "assembly_device.hpp"
#include "assembly_reduce.hpp"
#include "closure_device.hpp"
#include "continuation_kernels.hpp"
#include "flash_pipeline.hpp"

#include "../../include/mpmc_hnu_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <algorithm>
#include <vector>

namespace mpmc::dev {
MPMC_FOR_EACH_NC(MPMC_DECLARE_LAUNCH)
MPMC_FOR_EACH_NC(MPMC_DECLARE_PIPELINE)
MPMC_FOR_EACH_NC2(MPMC_DECLARE_CLOSURE)
MPMC_FOR_EACH_NC2(MPMC_DECLARE_ASSEMBLY)
} // namespace mpmc::dev

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& what) {
    g_last_error = what;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(MPMC_E_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

bool is_device_ptr(const void* ptr) {
    if (!ptr) { return false; }
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

} // namespace

namespace mpmc::api {
int record_error(int code, const std::string& what) { return fail(code, what); }
} // namespace mpmc::api

// Library-owned device residency of one prepared mixture.
struct mpmc_mixture {
    int nc;
    int device;
    double* dev_constants; // sqrt_tc, ac, b, kappa [nc each], factor [nc*nc], r_gas
    void* workspace;       // staged-schedule workspace used by mpmc_flash_batch
    size_t workspace_bytes;
    void* closure_ws;      // eligible-cell list of mpmc_closure_batch
    size_t closure_ws_bytes;
};

namespace {

size_t workspace_bytes_for(int nc, int64_t ncell) {
    switch (nc) {
#define MPMC_CASE(N) case N: return mpmc::dev::pipeline_bytes<N>(ncell);
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return 0;
    }
}

// Where the staged schedule's workspace comes from: the mixture's own (grown by
// hipMalloc), the caller's, or the static scratch of mpmc_flash_batch_static.
enum class Workspace { own, caller, static_scratch };

// Scratch of mpmc_flash_batch_static: the staged schedule's workspace for callers that
// cannot allocate. Every device that runs the library holds its own copy until the process
// ends. In host code `scratch` is only a stand-in for the symbol; the device address of
// each device's copy comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address reaches kernels.
static __device__ __align__(256) unsigned char scratch[MPMC_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address, looked up on the first call on that device.
int static_scratch(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { return hip_fail(e, "flash_batch_static: hipGetDevice"); }
    if (dev < 0 || dev >= kMaxDevices) { return fail(MPMC_E_DEVICE, "flash_batch_static: device index out of range"); }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            return fail(MPMC_E_DEVICE, std::string("flash_batch_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return MPMC_OK;
}

} // namespace

extern "C" {

const char* mpmc_last_error(void) { return g_last_error.c_str(); }

void mpmc_flash_default_options(mpmc_flash_options* o) {
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->fugacity_tolerance = 1e-11;
    o->max_phases = 3;
    o->precision = 0;
    o->schedule = 0;
}

int mpmc_mixture_create(int nc, const double* tc, const double* pc, const double* omega, const double* kij,
                        mpmc_mixture_t* out) {
    if (!out) { return fail(MPMC_E_INVALID_ARGUMENT, "mixture_create: null output handle"); }
    *out = nullptr;
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS) { return fail(MPMC_E_INVALID_ARGUMENT, "mixture_create: nc outside [1, 12]"); }
    if (!tc || !pc || !omega || !kij) { return fail(MPMC_E_INVALID_ARGUMENT, "mixture_create: null parameter array"); }
    // PrParameterSet / Pr76Pure validation (pr_parameters.hpp:81-89, pr76_pure.hpp:168-199)
    for (int i = 0; i < nc; ++i) {
        if (!std::isfinite(tc[i]) || !(tc[i] > 0) || !std::isfinite(pc[i]) || !(pc[i] > 0) || !std::isfinite(omega[i])) {
            return fail(MPMC_E_DOMAIN, "mixture_create: Tc, Pc must be finite and positive, omega finite");
        }
        for (int j = 0; j < nc; ++j) {
            const double k = kij[i * nc + j];
            if (!std::isfinite(k) || (i == j && k != 0.0) || k != kij[j * nc + i]) {
                return fail(MPMC_E_DOMAIN, "mixture_create: kij must be finite, symmetric, zero diagonal");
            }
        }
    }
    // Pr76Pure constructor arithmetic (pr76_pure.hpp:187-194), Pr76Mixture factors (pr76_mixture.hpp:236-250)
    const double r_gas = static_cast<double>(8.31446261815324L);
    std::vector<double> h(4 * nc + nc * nc + 1);
    double* sqrt_tc = h.data();
    double* ac = sqrt_tc + nc;
    double* b = ac + nc;
    double* kappa = b + nc;
    double* factor = kappa + nc;
    for (int i = 0; i < nc; ++i) {
        sqrt_tc[i] = std::sqrt(tc[i]);
        kappa[i] = static_cast<double>(0.37464L) +
                   omega[i] * (static_cast<double>(1.54226L) - static_cast<double>(0.26992L) * omega[i]);
        const double rtc = r_gas * tc[i];
        ac[i] = (static_cast<double>(0.45724L) * rtc) * (rtc / pc[i]);
        b[i] = (static_cast<double>(0.07780L) * rtc) / pc[i];
        if (!std::isfinite(kappa[i]) || !std::isfinite(ac[i]) || !(ac[i] > 0) || !std::isfinite(b[i]) || !(b[i] > 0)) {
            return fail(MPMC_E_DOMAIN, "mixture_create: precomputed PR76 coefficient out of range");
        }
    }
    for (int k = 0; k < nc * nc; ++k) { factor[k] = 1.0 - kij[k]; }
    h[4 * nc + nc * nc] = r_gas;
    auto* m = new (std::nothrow) mpmc_mixture{nc, 0, nullptr, nullptr, 0, nullptr, 0};
    if (!m) { return fail(MPMC_E_NO_MEMORY, "mixture_create: host allocation"); }
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) { e = hipMalloc(&m->dev_constants, h.size() * sizeof(double)); }
    if (e == hipSuccess) { e = hipMemcpy(m->dev_constants, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice); }
    if (e != hipSuccess) {
        if (m->dev_constants) { (void)hipFree(m->dev_constants); }
        delete m;
        return hip_fail(e, "mixture_create");
    }
    *out = m;
    return MPMC_OK;
}

int mpmc_mixture_destroy(mpmc_mixture_t m) {
    if (!m) { return MPMC_OK; }
    hipError_t e = m->dev_constants ? hipFree(m->dev_constants) : hipSuccess;
    for (void* ws : {m->workspace, m->closure_ws}) {
        if (ws) {
            const hipError_t e2 = hipFree(ws);
            if (e == hipSuccess) { e = e2; }
        }
    }
    delete m;
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "mixture_destroy");
}

int mpmc_mixture_components(mpmc_mixture_t m) { return m ? m->nc : 0; }

size_t mpmc_flash_workspace_bytes(int nc, int64_t ncell) {
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS || ncell < 0) { return 0; }
    return workspace_bytes_for(nc, ncell);
}

// Options of a flash call with defaults filled in; MPMC_OK or an error code.
static int flash_options(const mpmc_flash_options* options, mpmc_flash_options& o) {
    mpmc_flash_default_options(&o);
    if (options) { o = *options; }
    if (o.max_phases != 2 && o.max_phases != 3) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: max_phases must be 2 or 3");
    }
    if (!std::isfinite(o.fugacity_tolerance) || !(o.fugacity_tolerance > 0)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: fugacity tolerance must be finite and positive");
    }
    if (o.precision < 0 || o.precision > 2) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: precision must be 0, 1 or 2"); }
    if (o.schedule != 0 && o.schedule != 1) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: schedule must be 0 or 1"); }
    if (o.schedule == 1 && o.precision != 0) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: the serial schedule runs FP64 only");
    }
    if (o.group_mixtures < 0) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: group_mixtures must be >= 0"); }
    return MPMC_OK;
}

static int flash_buffers(const double* p, const double* t, const double* z, const mpmc_flash_outputs* out) {
    const void* ptrs[] = {p, t, z, out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                          out->ln_phi};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: buffers must be device memory"); }
    }
    if (out->counters && !is_device_ptr(out->counters)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: counters must be device memory");
    }
    return MPMC_OK;
}

// The workspace of a call: the mixture's own (grown to `need` bytes; the previous one is
// freed once the stream has drained) or the caller's, which must be device memory of at
// least `need` bytes.
static int acquire_workspace(mpmc_mixture_t m, size_t need, Workspace source, hipStream_t s, void*& workspace,
                             size_t& workspace_bytes, const char* caller_contract) {
    if (source == Workspace::own) {
        if (m->workspace_bytes < need) {
            if (m->workspace) {
                // the previous workspace may still be read by work queued on this stream
                const hipError_t e = hipStreamSynchronize(s);
                if (e != hipSuccess) { return hip_fail(e, "flash workspace sync"); }
                (void)hipFree(m->workspace);
                m->workspace = nullptr;
                m->workspace_bytes = 0;
            }
            if (hipMalloc(&m->workspace, need) != hipSuccess) {
                (void)hipGetLastError();
                m->workspace = nullptr;
                return fail(MPMC_E_NO_MEMORY, std::string("flash: workspace allocation of ") + std::to_string(need) +
                                                  " bytes failed");
            }
            m->workspace_bytes = need;
        }
        workspace = m->workspace;
        workspace_bytes = m->workspace_bytes;
        return MPMC_OK;
    }
    if (!workspace || workspace_bytes < need || !is_device_ptr(workspace)) {
        return fail(MPMC_E_INVALID_ARGUMENT, caller_contract);
    }
    return MPMC_OK;
}

static int flash_batch_impl(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                            const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                            size_t workspace_bytes, Workspace source, void* stream, int nhint = 0,
                            const double* hints = nullptr, int nfhint = 0, const double* fhints = nullptr) {
    if (!m || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: null mixture or outputs"); }
    if (ncell < 0 || ncell > (int64_t{1} << 31)) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: bad ncell"); }
    mpmc_flash_options o;
    if (const int rc = flash_options(options, o); rc != MPMC_OK) { return rc; }
    if (ncell == 0) { return MPMC_OK; }
    if (source == Workspace::static_scratch && o.schedule == 0 &&
        workspace_bytes_for(m->nc, ncell) > MPMC_STATIC_SCRATCH_BYTES) {
        return fail(MPMC_E_INVALID_ARGUMENT,
                    "flash_batch_static: " + std::to_string(ncell) + " cells need " +
                        std::to_string(workspace_bytes_for(m->nc, ncell)) + " bytes of scratch; the static scratch holds " +
                        std::to_string(static_cast<size_t>(MPMC_STATIC_SCRATCH_BYTES)) + " bytes");
    }
    if (const int rc = flash_buffers(p, t, z, out); rc != MPMC_OK) { return rc; }
    const mpmc::dev::FlashOut fo{out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                                 out->ln_phi, out->counters};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipErrorInvalidValue;
    if ((nhint > 0 || nfhint > 0) && o.schedule == 1) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_hinted: hints run on the staged schedule only");
    }
    if (o.schedule == 1) {
        switch (m->nc) {
#define MPMC_CASE(N)                                                                                                 \
    case N:                                                                                                          \
        e = mpmc::dev::launch_flash<N>(m->dev_constants, ncell, p, t, z, o.max_phases, o.fugacity_tolerance, fo, s); \
        break;
            MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
        default: return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: unsupported nc");
        }
        return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_batch launch");
    }
    const size_t need = workspace_bytes_for(m->nc, ncell);
    if (source == Workspace::static_scratch) {
        if (const int rc = static_scratch(&workspace); rc != MPMC_OK) { return rc; }
        workspace_bytes = MPMC_STATIC_SCRATCH_BYTES;
    } else if (const int rc = acquire_workspace(m, need, source, s, workspace, workspace_bytes,
                                                "flash_batch_ws: workspace must be device memory of "
                                                "mpmc_flash_workspace_bytes(nc, ncell) bytes");
               rc != MPMC_OK) {
        return rc;
    }
    switch (m->nc) {
#define MPMC_CASE(N)                                                                                             \
    case N:                                                                                                      \
        e = mpmc::dev::launch_pipeline<N>(m->dev_constants, 1, nullptr, workspace, workspace_bytes, ncell, p, t,  \
                                          z, o.max_phases, o.fugacity_tolerance, o.precision, fo, s, nhint,      \
                                          hints, nfhint, fhints, nullptr, nullptr);                              \
        break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_batch launch");
}

int mpmc_flash_batch(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                     const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* stream) {
    return flash_batch_impl(m, ncell, p, t, z, options, out, nullptr, 0, Workspace::own, stream);
}

int mpmc_flash_batch_ws(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                        const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                        size_t workspace_bytes, void* stream) {
    return flash_batch_impl(m, ncell, p, t, z, options, out, workspace, workspace_bytes, Workspace::caller, stream);
}

int mpmc_flash_batch_static(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                            const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* stream) {
    return flash_batch_impl(m, ncell, p, t, z, options, out, nullptr, 0, Workspace::static_scratch, stream);
}

int mpmc_flash_batch_warm(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                          int nhint, const double* hints, int nfinal, const double* final_hints,
                          const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                          size_t workspace_bytes, void* stream) {
    if (nhint < 0 || nhint > 6 || nfinal < 0 || nfinal > 6) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_warm: nhint and nfinal must be 0..6");
    }
    if ((nhint > 0 && ncell > 0 && !is_device_ptr(hints)) || (nfinal > 0 && ncell > 0 && !is_device_ptr(final_hints))) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_warm: hints must be device memory");
    }
    return flash_batch_impl(m, ncell, p, t, z, options, out, workspace, workspace_bytes,
                            workspace == nullptr ? Workspace::own : Workspace::caller, stream, nhint, hints, nfinal,
                            final_hints);
}

int mpmc_flash_batch_hinted(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                            int nhint, const double* hints, const mpmc_flash_options* options,
                            const mpmc_flash_outputs* out, void* workspace, size_t workspace_bytes, void* stream) {
    if (nhint < 0 || nhint > 6) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_hinted: nhint must be 0..6"); }
    if (nhint > 0 && ncell > 0 && !is_device_ptr(hints)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_hinted: hints must be device memory");
    }
    return flash_batch_impl(m, ncell, p, t, z, options, out, workspace, workspace_bytes,
                            workspace == nullptr ? Workspace::own : Workspace::caller, stream, nhint, hints);
}

// ---------------------------------------------------------------- ordered-path continuation
namespace {
constexpr int kContSlots = 6; // carried compositions (3) + fallback starts (3) per start list
// Continuation buffers after the pipeline workspace (continuation_kernels.hpp): feed
// [NC][n], initial and final-review starts [6][NC][n] each, three-phase start slots
// [2][3][NC][n] and seeds [2][2][n].
size_t a256(size_t v) { return (v + 255) & ~size_t{255}; }
size_t continuation_bytes(int nc, int64_t n) {
    const size_t N = static_cast<size_t>(n), NC = static_cast<size_t>(nc);
    return a256(workspace_bytes_for(nc, n)) + a256(NC * N * 8) + 2 * a256(kContSlots * NC * N * 8) +
           a256(6 * NC * N * 8) + a256(4 * N * 8);
}
} // namespace

void mpmc_continuation_default_options(mpmc_continuation_options* o) {
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->carry_compositions = 1;
    o->carry_three_phase = 1;
    o->start_seed[0] = 1.0 / 3.0;
    o->start_seed[1] = 1.0 / 3.0;
}

size_t mpmc_pt_continuation_workspace_bytes(int nc, int64_t npath) {
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS || npath < 0) { return 0; }
    return continuation_bytes(nc, npath);
}

int mpmc_pt_continuation(mpmc_mixture_t m, int64_t npath, int npoint, const double* p, const double* t,
                         const double* z, const mpmc_continuation_options* continuation,
                         const mpmc_flash_options* options, const mpmc_flash_outputs* out, int32_t* point_info,
                         void* workspace, size_t workspace_bytes, void* stream) {
    if (!m || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: null mixture or outputs"); }
    if (npath < 0 || npath > (int64_t{1} << 31)) { return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: bad npath"); }
    // Pr76PtContinuationOptions::max_points (pr76_pt_continuation.hpp:33, checked at :164-166)
    if (npoint < 1 || npoint > 4096) {
        return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: npoint must be 1..4096");
    }
    mpmc_continuation_options co;
    mpmc_continuation_default_options(&co);
    if (continuation) { co = *continuation; }
    if (co.nfallback_initial < 0 || co.nfallback_initial > 3 || co.nfallback_final < 0 || co.nfallback_final > 3 ||
        (co.three_phase_start != 0 && co.three_phase_start != 1)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: at most 3 fallback starts per list and 1 fixed "
                                             "three-phase start");
    }
    mpmc_flash_options o;
    if (const int rc = flash_options(options, o); rc != MPMC_OK) { return rc; }
    if (o.max_phases != 3) { return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: the path points are max3 flashes"); }
    if (o.schedule != 0) { return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: staged schedule only"); }
    if (npath == 0) { return MPMC_OK; }
    if (const int rc = flash_buffers(p, t, z, out); rc != MPMC_OK) { return rc; }
    if (point_info && !is_device_ptr(point_info)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: point_info must be device memory");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int nc = m->nc;
    const size_t need = continuation_bytes(nc, npath);
    if (const int rc = acquire_workspace(m, need, workspace == nullptr ? Workspace::own : Workspace::caller, s,
                                         workspace, workspace_bytes,
                                         "pt_continuation: workspace must be device memory of "
                                         "mpmc_pt_continuation_workspace_bytes(nc, npath) bytes");
        rc != MPMC_OK) {
        return rc;
    }
    const size_t pipe_bytes = a256(workspace_bytes_for(nc, npath));
    const size_t N = static_cast<size_t>(npath), NC = static_cast<size_t>(nc);
    char* at = static_cast<char*>(workspace) + pipe_bytes;
    auto carve = [&](size_t bytes) {
        double* ptr = reinterpret_cast<double*>(at);
        at += a256(bytes);
        return ptr;
    };
    double* feed = carve(NC * N * 8);
    double* hints = carve(kContSlots * NC * N * 8);
    double* fhints = carve(kContSlots * NC * N * 8);
    double* tpz = carve(6 * NC * N * 8);
    double* tpf = carve(4 * N * 8);
    mpmc::dev::ContStarts cs{};
    cs.carry = (co.carry_compositions ? 1 : 0) | (co.carry_three_phase ? 2 : 0);
    cs.nfi = co.nfallback_initial;
    cs.nff = co.nfallback_final;
    cs.has_start = co.three_phase_start;
    for (int k = 0; k < 3; ++k) {
        for (int i = 0; i < nc; ++i) {
            cs.fi[k][i] = co.fallback_initial[k][i];
            cs.ff[k][i] = co.fallback_final[k][i];
            cs.st[k][i] = co.start[k][i];
        }
    }
    cs.sf[0] = co.start_seed[0];
    cs.sf[1] = co.start_seed[1];
    const int carried = (cs.carry & 1) ? 3 : 0;
    const int nh = carried + cs.nfi, nfh = carried + cs.nff;
    const bool slots3 = (cs.carry & 2) || cs.has_start;
    const mpmc::dev::FlashOut fo{out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                                 out->ln_phi, out->counters};
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((npath + 255) / 256, 2048));
    hipError_t e = hipSuccess;
    switch (nc) {
#define MPMC_CASE(N)                                                                                                  \
    case N:                                                                                                           \
        hipLaunchKernelGGL(mpmc::dev::k_cont_feed<N>, dim3(grid), dim3(256), 0, s, npath, z, feed);                   \
        for (int k = 0; k < npoint && e == hipSuccess; ++k) {                                                         \
            const int64_t off = static_cast<int64_t>(k) * npath;                                                      \
            hipLaunchKernelGGL(mpmc::dev::k_cont_hints<N>, dim3(grid), dim3(256), 0, s, npath,                        \
                               mpmc::dev::point_slice<N>(fo, npath, k > 0 ? k - 1 : 0), k > 0 ? 1 : 0, feed, cs, nh,  \
                               nfh, hints, fhints, slots3 ? tpz : nullptr, tpf, point_info ? point_info + off : nullptr); \
            e = hipGetLastError();                                                                                    \
            if (e == hipSuccess) {                                                                                    \
                e = mpmc::dev::launch_pipeline<N>(m->dev_constants, 1, nullptr, workspace, pipe_bytes, npath, p + off, \
                                                  t + off, feed, 3, o.fugacity_tolerance, o.precision,                \
                                                  mpmc::dev::point_slice<N>(fo, npath, k), s, nh, hints, nfh, fhints, \
                                                  slots3 ? tpz : nullptr, slots3 ? tpf : nullptr);                    \
            }                                                                                                         \
        }                                                                                                             \
        if (e == hipSuccess) {                                                                                        \
            hipLaunchKernelGGL(mpmc::dev::k_cont_reject<N>, dim3(grid), dim3(256), 0, s, npath, npoint, fo,           \
                               point_info);                                                                           \
            e = hipGetLastError();                                                                                    \
        }                                                                                                             \
        break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "pt_continuation: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "pt_continuation launch");
}

// ---------------------------------------------------------------- ragged multi-mixture batch
namespace {
constexpr size_t kMixTableLds = 48 * 1024; // largest mixture table one launch stages in LDS
size_t mix_doubles(int nc) { return static_cast<size_t>(4 * nc + nc * nc + 1); }
size_t align256(size_t v) { return (v + 255) & ~size_t{255}; }
} // namespace

static __global__ void k_fill_mix(int32_t* mix_of, int64_t n, int32_t value) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        mix_of[i] = value;
    }
}

int mpmc_flash_multi_max_mixtures(int nc) {
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS) { return 0; }
    return static_cast<int>(kMixTableLds / (mix_doubles(nc) * sizeof(double)));
}

size_t mpmc_flash_multi_workspace_bytes(int nc, int nmix, int64_t ncell) {
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS || nmix < 1 || ncell < 0) { return 0; }
    return align256(workspace_bytes_for(nc, ncell)) + align256(static_cast<size_t>(nmix) * mix_doubles(nc) * 8) +
           align256(static_cast<size_t>(ncell) * sizeof(int32_t));
}

int mpmc_flash_multi(int nmix, const mpmc_mixture_t* mixtures, const int64_t* cell_offset, const double* p,
                     const double* t, const double* z, const mpmc_flash_options* options,
                     const mpmc_flash_outputs* out, void* workspace, size_t workspace_bytes, void* stream) {
    if (nmix < 1 || !mixtures || !cell_offset || !out) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: null argument or no mixture");
    }
    const int nc = mixtures[0] ? mixtures[0]->nc : 0;
    for (int k = 0; k < nmix; ++k) {
        if (!mixtures[k]) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: null mixture"); }
        if (mixtures[k]->nc != nc) {
            return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: all mixtures of one call must have the same nc");
        }
    }
    if (nmix > mpmc_flash_multi_max_mixtures(nc)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: more mixtures than mpmc_flash_multi_max_mixtures(nc)");
    }
    if (cell_offset[0] != 0) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: cell_offset[0] must be 0"); }
    for (int k = 0; k < nmix; ++k) {
        if (cell_offset[k + 1] < cell_offset[k]) {
            return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: cell offsets must not decrease");
        }
    }
    const int64_t ncell = cell_offset[nmix];
    if (ncell > (int64_t{1} << 31)) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: bad ncell"); }
    mpmc_flash_options o;
    if (const int rc = flash_options(options, o); rc != MPMC_OK) { return rc; }
    if (o.schedule != 0) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: runs the staged schedule only"); }
    if (ncell == 0) { return MPMC_OK; }
    if (const int rc = flash_buffers(p, t, z, out); rc != MPMC_OK) { return rc; }
    const size_t need = mpmc_flash_multi_workspace_bytes(nc, nmix, ncell);
    if (!workspace || workspace_bytes < need || !is_device_ptr(workspace)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: workspace must be device memory of "
                                             "mpmc_flash_multi_workspace_bytes(nc, nmix, ncell) bytes");
    }
    char* base = static_cast<char*>(workspace);
    const size_t pipe_bytes = align256(workspace_bytes_for(nc, ncell));
    double* table = reinterpret_cast<double*>(base + pipe_bytes);
    int32_t* mix_of = reinterpret_cast<int32_t*>(base + pipe_bytes +
                                                 align256(static_cast<size_t>(nmix) * mix_doubles(nc) * 8));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    for (int k = 0; k < nmix && e == hipSuccess; ++k) {
        e = hipMemcpyAsync(table + k * mix_doubles(nc), mixtures[k]->dev_constants, mix_doubles(nc) * sizeof(double),
                           hipMemcpyDeviceToDevice, s);
        const int64_t len = cell_offset[k + 1] - cell_offset[k];
        if (e == hipSuccess && len > 0) {
            const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((len + 255) / 256, 4096));
            hipLaunchKernelGGL(k_fill_mix, dim3(blocks), dim3(256), 0, s, mix_of + cell_offset[k], len, k);
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) { return hip_fail(e, "flash_multi setup"); }
    const mpmc::dev::FlashOut fo{out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                                 out->ln_phi, out->counters};
    e = hipErrorInvalidValue;
    switch (nc) {
#define MPMC_CASE(N)                                                                                             \
    case N:                                                                                                      \
        e = mpmc::dev::launch_pipeline<N>(table, nmix, nmix > 1 ? mix_of : nullptr, workspace, pipe_bytes, ncell, \
                                          p, t, z, o.max_phases, o.fugacity_tolerance, o.precision, fo, s, 0,   \
                                          nullptr, 0, nullptr, nullptr, nullptr);                               \
        break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "flash_multi: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_multi launch");
}

// ---------------------------------------------------------------- flash plans
} // extern "C"

namespace {

// One strided copy: rows x n elements of `bytes` each, from src (row pitch src_pitch
// elements) to dst (row pitch dst_pitch elements).
struct SegCopy {
    const void* src;
    void* dst;
    int64_t n, src_pitch, dst_pitch;
    int32_t rows, bytes;
};

__global__ void k_seg_copy(const SegCopy* __restrict__ list) {
    const SegCopy d = list[blockIdx.y];
    const int64_t total = d.n * d.rows;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = e / d.n, i = e % d.n;
        if (d.bytes == 8) {
            static_cast<double*>(d.dst)[r * d.dst_pitch + i] = static_cast<const double*>(d.src)[r * d.src_pitch + i];
        } else {
            static_cast<int32_t*>(d.dst)[r * d.dst_pitch + i] = static_cast<const int32_t*>(d.src)[r * d.src_pitch + i];
        }
    }
}

// A launch group: segments of one nc whose mixtures fit one LDS table, packed into one
// ragged batch of n cells.
struct PlanGroup {
    int nc = 0;
    int64_t n = 0;
    std::vector<mpmc_mixture_t> mixtures;  // table order
    std::vector<int64_t> offsets;          // per table entry: first cell (cells are grouped by mixture)
    double *p = nullptr, *t = nullptr, *z = nullptr;
    mpmc_flash_outputs out{};
    double* table = nullptr;
    int32_t* mix_of = nullptr;
    SegCopy* gather = nullptr;             // device descriptor lists
    SegCopy* scatter = nullptr;
    int ngather = 0, nscatter = 0;
};

} // namespace

struct mpmc_flash_plan {
    mpmc_flash_options options{};
    std::vector<PlanGroup> groups;
    std::vector<void*> allocations;
    void* workspace = nullptr;
    size_t workspace_bytes = 0;
    size_t device_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

hipError_t plan_alloc(mpmc_flash_plan* pl, void** ptr, size_t bytes) {
    *ptr = nullptr;
    const hipError_t e = hipMalloc(ptr, bytes > 0 ? bytes : 1);
    if (e == hipSuccess) {
        pl->allocations.push_back(*ptr);
        pl->device_bytes += bytes;
    }
    return e;
}

// The launch sequence of every group on stream s: gather the segments' inputs, run the
// ragged pipeline, scatter the outputs back to the segments.
hipError_t plan_record(mpmc_flash_plan* pl, hipStream_t s) {
    for (PlanGroup& g : pl->groups) {
        hipLaunchKernelGGL(k_seg_copy, dim3(256, static_cast<unsigned>(g.ngather)), dim3(256), 0, s, g.gather);
        const mpmc::dev::FlashOut fo{g.out.status, g.out.phase_count, g.out.beta, g.out.zfactor, g.out.branch,
                                     g.out.x, g.out.ln_phi, g.out.counters};
        const int nmix = static_cast<int>(g.mixtures.size());
        hipError_t e = hipErrorInvalidValue;
        switch (g.nc) {
#define MPMC_CASE(N)                                                                                              \
    case N:                                                                                                       \
        e = mpmc::dev::launch_pipeline<N>(g.table, nmix, nmix > 1 ? g.mix_of : nullptr, pl->workspace,            \
                                          pl->workspace_bytes, g.n, g.p, g.t, g.z, pl->options.max_phases,        \
                                          pl->options.fugacity_tolerance, pl->options.precision, fo, s, 0,        \
                                          nullptr, 0, nullptr, nullptr, nullptr);                                 \
        break;
            MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
        default: break;
        }
        if (e != hipSuccess) { return e; }
        hipLaunchKernelGGL(k_seg_copy, dim3(256, static_cast<unsigned>(g.nscatter)), dim3(256), 0, s, g.scatter);
    }
    return hipGetLastError();
}

} // namespace

extern "C" {

int mpmc_flash_plan_create(int nseg, const mpmc_flash_segment* segs, const mpmc_flash_options* options,
                           mpmc_flash_plan_t* out) {
    if (!out) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_create: null output handle"); }
    *out = nullptr;
    if (nseg < 1 || !segs) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_create: no segments"); }
    mpmc_flash_options o;
    if (const int rc = flash_options(options, o); rc != MPMC_OK) { return rc; }
    if (o.schedule != 0) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_create: staged schedule only"); }
    const bool counters = segs[0].out.counters != nullptr;
    for (int k = 0; k < nseg; ++k) {
        const mpmc_flash_segment& g = segs[k];
        if (!g.mixture || g.ncell < 0 || g.ncell > (int64_t{1} << 30)) {
            return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_create: bad segment");
        }
        if (g.ncell > 0) {
            if (const int rc = flash_buffers(g.p, g.t, g.z, &g.out); rc != MPMC_OK) { return rc; }
        }
        if ((g.out.counters != nullptr) != counters) {
            return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_create: counters must be given for all segments or none");
        }
    }
    auto pl = std::make_unique<mpmc_flash_plan>();
    pl->options = o;
    // segments by component count, then in the given order; a group closes when its
    // mixture table is full
    std::vector<int> order(nseg);
    for (int k = 0; k < nseg; ++k) { order[k] = k; }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return segs[a].mixture->nc < segs[b].mixture->nc; });
    std::vector<std::vector<int>> members;
    for (int k : order) {
        if (segs[k].ncell == 0) { continue; }
        const int nc = segs[k].mixture->nc;
        bool fits = !pl->groups.empty() && pl->groups.back().nc == nc;
        if (fits) {
            PlanGroup& g = pl->groups.back();
            const bool known = std::find(g.mixtures.begin(), g.mixtures.end(), segs[k].mixture) != g.mixtures.end();
            const int cap = o.group_mixtures > 0 ? std::min(o.group_mixtures, mpmc_flash_multi_max_mixtures(nc))
                                                 : mpmc_flash_multi_max_mixtures(nc);
            fits = known || static_cast<int>(g.mixtures.size()) < cap;
            fits = fits && g.n + segs[k].ncell <= (int64_t{1} << 31);
        }
        if (!fits) {
            pl->groups.emplace_back();
            pl->groups.back().nc = nc;
            members.emplace_back();
        }
        PlanGroup& g = pl->groups.back();
        if (std::find(g.mixtures.begin(), g.mixtures.end(), segs[k].mixture) == g.mixtures.end()) {
            g.mixtures.push_back(segs[k].mixture);
        }
        g.n += segs[k].ncell;
        members.back().push_back(k);
    }
    if (pl->groups.empty()) {
        *out = pl.release();
        return MPMC_OK;
    }
    hipError_t e = hipSuccess;
    size_t ws_need = 0;
    for (size_t gi = 0; gi < pl->groups.size() && e == hipSuccess; ++gi) {
        PlanGroup& g = pl->groups[gi];
        // cells of the group ordered by table entry, segments in the given order within one
        const int nmix = static_cast<int>(g.mixtures.size());
        std::vector<std::vector<int>> by_mix(nmix);
        for (int k : members[gi]) {
            const int slot = static_cast<int>(std::find(g.mixtures.begin(), g.mixtures.end(), segs[k].mixture) -
                                              g.mixtures.begin());
            by_mix[slot].push_back(k);
        }
        const int64_t n = g.n;
        const int nc = g.nc;
        void* ptr = nullptr;
        auto take = [&](auto*& dst, size_t count) {
            using T = std::remove_reference_t<decltype(*dst)>;
            if (e == hipSuccess) { e = plan_alloc(pl.get(), &ptr, count * sizeof(T)); }
            dst = static_cast<T*>(e == hipSuccess ? ptr : nullptr);
        };
        take(g.p, n);
        take(g.t, n);
        take(g.z, static_cast<size_t>(nc) * n);
        take(g.out.status, n);
        take(g.out.phase_count, n);
        take(g.out.beta, 3 * n);
        take(g.out.zfactor, 3 * n);
        take(g.out.branch, 3 * n);
        take(g.out.x, static_cast<size_t>(3 * nc) * n);
        take(g.out.ln_phi, static_cast<size_t>(3 * nc) * n);
        if (counters) { take(g.out.counters, 4 * n); }
        take(g.table, static_cast<size_t>(nmix) * mix_doubles(nc));
        take(g.mix_of, n);
        std::vector<SegCopy> gather, scatter;
        int64_t at = 0;
        g.offsets.assign(nmix, 0);
        for (int slot = 0; slot < nmix && e == hipSuccess; ++slot) {
            g.offsets[slot] = at;
            const int64_t first = at;
            for (int k : by_mix[slot]) {
                const mpmc_flash_segment& sg = segs[k];
                const int64_t m = sg.ncell;
                gather.push_back({sg.p, g.p + at, m, m, n, 1, 8});
                gather.push_back({sg.t, g.t + at, m, m, n, 1, 8});
                gather.push_back({sg.z, g.z + at, m, m, n, nc, 8});
                scatter.push_back({g.out.status + at, sg.out.status, m, n, m, 1, 4});
                scatter.push_back({g.out.phase_count + at, sg.out.phase_count, m, n, m, 1, 4});
                scatter.push_back({g.out.beta + at, sg.out.beta, m, n, m, 3, 8});
                scatter.push_back({g.out.zfactor + at, sg.out.zfactor, m, n, m, 3, 8});
                scatter.push_back({g.out.branch + at, sg.out.branch, m, n, m, 3, 4});
                scatter.push_back({g.out.x + at, sg.out.x, m, n, m, 3 * nc, 8});
                scatter.push_back({g.out.ln_phi + at, sg.out.ln_phi, m, n, m, 3 * nc, 8});
                if (counters) { scatter.push_back({g.out.counters + at, sg.out.counters, m, n, m, 4, 4}); }
                at += m;
            }
            e = hipMemcpy(g.table + slot * mix_doubles(nc), g.mixtures[slot]->dev_constants,
                          mix_doubles(nc) * sizeof(double), hipMemcpyDeviceToDevice);
            if (e == hipSuccess && at > first) {
                hipLaunchKernelGGL(k_fill_mix, dim3(static_cast<unsigned>(std::min<int64_t>((at - first + 255) / 256, 4096))),
                                   dim3(256), 0, nullptr, g.mix_of + first, at - first, slot);
                e = hipGetLastError();
            }
        }
        g.ngather = static_cast<int>(gather.size());
        g.nscatter = static_cast<int>(scatter.size());
        take(g.gather, gather.size());
        take(g.scatter, scatter.size());
        if (e == hipSuccess) { e = hipMemcpy(g.gather, gather.data(), gather.size() * sizeof(SegCopy), hipMemcpyHostToDevice); }
        if (e == hipSuccess) { e = hipMemcpy(g.scatter, scatter.data(), scatter.size() * sizeof(SegCopy), hipMemcpyHostToDevice); }
        ws_need = std::max(ws_need, workspace_bytes_for(nc, n));
    }
    if (e == hipSuccess) {
        e = plan_alloc(pl.get(), &pl->workspace, ws_need);
        pl->workspace_bytes = ws_need;
    }
    if (e == hipSuccess) { e = hipDeviceSynchronize(); }
    if (e != hipSuccess) {
        mpmc_flash_plan_destroy(pl.release());
        return e == hipErrorOutOfMemory ? fail(MPMC_E_NO_MEMORY, "flash_plan_create: device allocation failed")
                                        : hip_fail(e, "flash_plan_create");
    }
    // One eager run settles every lazily computed launch parameter, then the sequence is
    // captured once into a graph. The eager run goes to the null stream, not to the capture
    // stream: a capture stream only records, so its hardware queue never has to back the
    // pipeline's large private segments (the serial fallback and the three-phase tail carry
    // up to 22 KB per lane), which the runtime has refused on an additional queue while
    // another queue held its scratch.
    hipStream_t cs = nullptr;
    e = plan_record(pl.get(), nullptr);
    if (e == hipSuccess) { e = hipStreamSynchronize(nullptr); }
    if (e == hipSuccess) { e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking); }
    if (e == hipSuccess) { e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal); }
    if (e == hipSuccess) {
        const hipError_t rec = plan_record(pl.get(), cs);
        e = hipStreamEndCapture(cs, &pl->graph);
        if (rec != hipSuccess) { e = rec; }
    }
    if (e == hipSuccess) { e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0); }
    if (cs) { (void)hipStreamDestroy(cs); }
    if (e != hipSuccess) {
        mpmc_flash_plan_destroy(pl.release());
        return hip_fail(e, "flash_plan_create capture");
    }
    *out = pl.release();
    return MPMC_OK;
}

int mpmc_flash_plan_launch(mpmc_flash_plan_t pl, void* stream) {
    if (!pl) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_launch: null plan"); }
    if (!pl->exec) { return MPMC_OK; }
    const hipError_t e = hipGraphLaunch(pl->exec, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_plan_launch");
}

int mpmc_flash_plan_info(mpmc_flash_plan_t pl, int32_t* groups, size_t* device_bytes) {
    if (!pl) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_plan_info: null plan"); }
    if (groups) { *groups = static_cast<int32_t>(pl->groups.size()); }
    if (device_bytes) { *device_bytes = pl->device_bytes; }
    return MPMC_OK;
}

int mpmc_flash_plan_destroy(mpmc_flash_plan_t pl) {
    if (!pl) { return MPMC_OK; }
    hipError_t e = hipSuccess;
    if (pl->exec || !pl->allocations.empty()) { e = hipDeviceSynchronize(); }
    if (pl->exec) { (void)hipGraphExecDestroy(pl->exec); }
    if (pl->graph) { (void)hipGraphDestroy(pl->graph); }
    for (void* ptr : pl->allocations) { (void)hipFree(ptr); }
    delete pl;
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_plan_destroy");
}

int mpmc_phase_batch(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* x,
                     int32_t* root_status, int32_t* root_count, double* zroot, int32_t* root_flags, double* ln_phi,
                     void* stream) {
    if (!m) { return fail(MPMC_E_INVALID_ARGUMENT, "phase_batch: null mixture"); }
    if (ncell < 0 || ncell > (int64_t{1} << 40)) { return fail(MPMC_E_INVALID_ARGUMENT, "phase_batch: bad ncell"); }
    if (ncell == 0) { return MPMC_OK; }
    const void* ptrs[] = {p, t, x, root_status, root_count, zroot, root_flags, ln_phi};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "phase_batch: buffers must be device memory"); }
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipErrorInvalidValue;
    switch (m->nc) {
#define MPMC_CASE(N)                                                                                                   \
    case N:                                                                                                            \
        e = mpmc::dev::launch_phase<N>(m->dev_constants, ncell, p, t, x, root_status, root_count, zroot, root_flags, \
                                       ln_phi, s);                                                                     \
        break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "phase_batch: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "phase_batch launch");
}

int mpmc_rr_batch(int nc, int64_t n, const double* z, const double* log_k, const double* guess, int32_t* status,
                  double* beta, double* x, double* y, int32_t* iterations, void* stream) {
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS || n < 0 || n > (int64_t{1} << 31)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "rr_batch: nc outside [1, 12] or bad n");
    }
    if (n == 0) { return MPMC_OK; }
    const void* ptrs[] = {z, log_k, status, beta, x, y, iterations};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "rr_batch: buffers must be device memory"); }
    }
    if (guess && !is_device_ptr(guess)) { return fail(MPMC_E_INVALID_ARGUMENT, "rr_batch: guess must be device memory"); }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipErrorInvalidValue;
    switch (nc) {
#define MPMC_CASE(N) case N: e = mpmc::dev::launch_rr<N>(n, z, log_k, guess, status, beta, x, y, iterations, s); break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: break;
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "rr_batch launch");
}

void mpmc_closure_default_options(mpmc_closure_options* o) {
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->minimum_derivative_phase_fraction = 1.490116119384765625e-8;
    o->minimum_reciprocal_condition = 0.0;
    o->fugacity_tolerance = 1e-11;
}

int mpmc_closure_batch(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                       const mpmc_flash_outputs* flash, const mpmc_closure_options* options,
                       const mpmc_closure_outputs* out, void* stream) {
    if (!m || !flash || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: null mixture, flash or outputs"); }
    if (m->nc < 2) { return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: the closure needs nc >= 2"); }
    if (ncell < 0 || ncell > (int64_t{1} << 31)) { return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: bad ncell"); }
    mpmc_closure_options o;
    mpmc_closure_default_options(&o);
    if (options) { o = *options; }
    // PtSensitivityOptions validation (pr76_sensitivity.hpp:56-65)
    if (!std::isfinite(o.minimum_derivative_phase_fraction) || !(o.minimum_derivative_phase_fraction > 0.0) ||
        !(o.minimum_derivative_phase_fraction < 0.5) || !std::isfinite(o.minimum_reciprocal_condition) ||
        o.minimum_reciprocal_condition < 0.0 || o.minimum_reciprocal_condition >= 1.0) {
        return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: invalid sensitivity options");
    }
    if (!std::isfinite(o.fugacity_tolerance) || !(o.fugacity_tolerance > 0)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: fugacity tolerance must be finite and positive");
    }
    if (ncell == 0) { return MPMC_OK; }
    const void* ptrs[] = {p, t, z, flash->status, flash->phase_count, flash->beta, flash->zfactor, flash->branch,
                          flash->x, out->status, out->primal, out->jac};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: buffers must be device memory"); }
    }
    if ((out->log_k_jac && !is_device_ptr(out->log_k_jac)) || (out->diag && !is_device_ptr(out->diag))) {
        return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: optional outputs must be device memory");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t need = mpmc::dev::closure_workspace_bytes(ncell);
    hipError_t e = hipSuccess;
    if (m->closure_ws_bytes < need) {
        if (m->closure_ws) {
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) { return hip_fail(e, "closure_batch workspace sync"); }
            (void)hipFree(m->closure_ws);
            m->closure_ws = nullptr;
            m->closure_ws_bytes = 0;
        }
        e = hipMalloc(&m->closure_ws, need);
        if (e != hipSuccess) {
            m->closure_ws = nullptr;
            return fail(MPMC_E_NO_MEMORY, "closure_batch: workspace allocation failed");
        }
        m->closure_ws_bytes = need;
    }
    const mpmc::dev::ClosureIn in{flash->status, flash->phase_count, flash->beta, flash->zfactor, flash->branch,
                                  flash->x};
    const mpmc::dev::ClosureOut co{out->status, out->primal, out->jac, out->log_k_jac, out->diag};
    const mpmc::dev::ClosureOpt opt{o.minimum_derivative_phase_fraction, o.minimum_reciprocal_condition,
                                    o.fugacity_tolerance};
    e = hipErrorInvalidValue;
    switch (m->nc) {
#define MPMC_CASE(N)                                                                                            \
    case N:                                                                                                     \
        e = mpmc::dev::launch_closure<N>(m->dev_constants, m->closure_ws, m->closure_ws_bytes, ncell, p, t, z, \
                                         in, co, opt, s);                                                       \
        break;
        MPMC_FOR_EACH_NC2(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "closure_batch: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "closure_batch launch");
}

// ---------------------------------------------------------------- two-phase assembly
struct mpmc_tpfa_grid {
    int64_t ncell, nface;
    const double *bulk, *poro, *trans, *gdot, *gth;
    const int32_t *owner, *neighbour;
    int32_t* adj_off = nullptr;  // device CSR cell -> faces
    int32_t* adj_face = nullptr;
    int32_t* slot_face = nullptr; // device [max_degree][ncell], -1 past a cell's degree
    int32_t* slot_other = nullptr; // device [max_degree][ncell], cell across (+ 2^30 as neighbour), -1 past
    int32_t max_degree = 0;
    int32_t block_stride = 1;     // largest face span |owner - neighbour| in 64-cell blocks
    void* rec = nullptr;         // per-cell face records
    size_t rec_bytes = 0;
    unsigned long long* flag = nullptr;
};

int mpmc_tpfa_grid_create(int64_t ncell, int64_t nface, const double* bulk, const double* poro, const int32_t* owner,
                          const int32_t* neighbour, const double* trans, const double* gdot, const double* gth,
                          mpmc_tpfa_grid_t* out) {
    if (!out) { return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_create: null handle"); }
    *out = nullptr;
    if (ncell < 1 || nface < 0 || ncell > (int64_t{1} << 30) || nface > (int64_t{1} << 30)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_create: bad sizes");
    }
    const void* ptrs[] = {bulk, poro, trans, gdot, gth, owner, neighbour};
    for (const void* ptr : ptrs) {
        if (nface == 0 && (ptr == trans || ptr == gdot || ptr == gth || ptr == owner || ptr == neighbour)) { continue; }
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_create: arrays must be device memory"); }
    }
    // cell -> face adjacency, faces of each cell in face order (host counting sort, once per grid)
    std::vector<int32_t> ho(static_cast<size_t>(nface)), hn(static_cast<size_t>(nface));
    hipError_t e = hipSuccess;
    if (nface > 0) {
        e = hipMemcpy(ho.data(), owner, sizeof(int32_t) * nface, hipMemcpyDeviceToHost);
        if (e == hipSuccess) { e = hipMemcpy(hn.data(), neighbour, sizeof(int32_t) * nface, hipMemcpyDeviceToHost); }
        if (e != hipSuccess) { return hip_fail(e, "tpfa_grid_create"); }
    }
    std::vector<int32_t> off(static_cast<size_t>(ncell) + 1, 0), adj(static_cast<size_t>(2 * nface));
    for (int64_t f = 0; f < nface; ++f) {
        if (ho[f] < 0 || ho[f] >= ncell || hn[f] < 0 || hn[f] >= ncell || ho[f] == hn[f]) {
            return fail(MPMC_E_DOMAIN, "tpfa_grid_create: face cells out of range");
        }
        ++off[ho[f] + 1];
        ++off[hn[f] + 1];
    }
    for (int64_t c = 0; c < ncell; ++c) { off[c + 1] += off[c]; }
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t f = 0; f < nface; ++f) {
        adj[fill[ho[f]]++] = static_cast<int32_t>(f);
        adj[fill[hn[f]]++] = static_cast<int32_t>(f);
    }
    // the same adjacency by (slot, cell), cell fastest: the layout of the off-diagonal blocks
    int32_t maxdeg = 0;
    for (int64_t c = 0; c < ncell; ++c) { maxdeg = std::max(maxdeg, off[c + 1] - off[c]); }
    int64_t span = 0;
    for (int64_t f = 0; f < nface; ++f) { span = std::max<int64_t>(span, std::abs(int64_t{ho[f]} - int64_t{hn[f]})); }
    std::vector<int32_t> slots(static_cast<size_t>(maxdeg) * static_cast<size_t>(ncell), -1);
    std::vector<int32_t> others(slots.size(), -1);
    for (int64_t c = 0; c < ncell; ++c) {
        for (int32_t a = off[c]; a < off[c + 1]; ++a) {
            const size_t k = static_cast<size_t>(a - off[c]) * ncell + c;
            const int32_t f = adj[a];
            slots[k] = f;
            others[k] = ho[f] == c ? hn[f] : ho[f] + mpmc::dev::kSlotNeighbour;
        }
    }
    auto* g = new (std::nothrow) mpmc_tpfa_grid{ncell, nface, bulk, poro, trans, gdot, gth, owner, neighbour};
    if (!g) { return fail(MPMC_E_NO_MEMORY, "tpfa_grid_create: host allocation"); }
    g->max_degree = maxdeg;
    g->block_stride = static_cast<int32_t>(std::max<int64_t>(1, span / 64));
    e = hipMalloc(&g->adj_off, sizeof(int32_t) * off.size());
    if (e == hipSuccess) { e = hipMalloc(&g->adj_face, sizeof(int32_t) * std::max<size_t>(1, adj.size())); }
    if (e == hipSuccess) { e = hipMalloc(&g->slot_face, sizeof(int32_t) * std::max<size_t>(1, slots.size())); }
    if (e == hipSuccess) { e = hipMalloc(&g->slot_other, sizeof(int32_t) * std::max<size_t>(1, others.size())); }
    if (e == hipSuccess) { e = hipMalloc(&g->flag, 256); }
    if (e == hipSuccess) { e = hipMemcpy(g->adj_off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice); }
    if (e == hipSuccess && !adj.empty()) {
        e = hipMemcpy(g->adj_face, adj.data(), sizeof(int32_t) * adj.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !slots.empty()) {
        e = hipMemcpy(g->slot_face, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !others.empty()) {
        e = hipMemcpy(g->slot_other, others.data(), sizeof(int32_t) * others.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        mpmc_tpfa_grid_destroy(g);
        return hip_fail(e, "tpfa_grid_create");
    }
    *out = g;
    return MPMC_OK;
}

int mpmc_tpfa_grid_destroy(mpmc_tpfa_grid_t g) {
    if (!g) { return MPMC_OK; }
    for (void* ptr : {static_cast<void*>(g->adj_off), static_cast<void*>(g->adj_face),
                      static_cast<void*>(g->slot_face), static_cast<void*>(g->slot_other), g->rec,
                      static_cast<void*>(g->flag)}) {
        if (ptr) { (void)hipFree(ptr); }
    }
    delete g;
    return MPMC_OK;
}

int mpmc_tpfa_grid_adjacency(mpmc_tpfa_grid_t g, const int32_t** adj_off, const int32_t** adj_face) {
    if (!g || !adj_off || !adj_face) { return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_adjacency: null argument"); }
    *adj_off = g->adj_off;
    *adj_face = g->adj_face;
    return MPMC_OK;
}

int mpmc_tpfa_grid_cells(mpmc_tpfa_grid_t g, int64_t* ncell) {
    if (!g || !ncell) { return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_cells: null argument"); }
    *ncell = g->ncell;
    return MPMC_OK;
}

int mpmc_tpfa_grid_max_degree(mpmc_tpfa_grid_t g, int32_t* max_degree) {
    if (!g || !max_degree) { return fail(MPMC_E_INVALID_ARGUMENT, "tpfa_grid_max_degree: null argument"); }
    *max_degree = g->max_degree;
    return MPMC_OK;
}

void mpmc_assembly_default_model(int nc, mpmc_assembly_model* md) {
    if (!md) { return; }
    std::memset(md, 0, sizeof(*md));
    md->mu_a0 = 1.5e-5;
    md->mu_a1 = 1.0e-9;
    md->cp = 2000.0;
    md->t0 = 273.15;
    md->cv_rock = 2.0e6;
    for (int i = 0; i < nc && i < MPMC_MAX_COMPONENTS; ++i) { md->molar_mass[i] = 0.016 + 0.014 * i; }
}

// Argument contract of one Newton evaluation; on success `need` is the record workspace size.
static int assembly_check(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md, const double* q,
                          const int32_t* branch, const double* prev, double dt, const mpmc_assembly_outputs* out,
                          bool inventory, size_t& need) {
    if (!m || !g || !md || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: null argument"); }
    const int nc = m->nc;
    if (nc < 2) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: two-phase assembly needs nc >= 2"); }
    if (!inventory && (!std::isfinite(dt) || !(dt > 0))) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: dt must be positive"); }
    for (int i = 0; i < 5; ++i) {
        const double v = (&md->mu_a0)[i];
        if (!std::isfinite(v)) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: model parameters must be finite"); }
    }
    for (int i = 0; i < nc; ++i) {
        if (!std::isfinite(md->molar_mass[i]) || !(md->molar_mass[i] > 0)) {
            return fail(MPMC_E_INVALID_ARGUMENT, "assembly: molar masses must be positive");
        }
    }
    const void* need_ptrs[] = {q, branch, out->cell_status, out->residual};
    for (const void* ptr : need_ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: buffers must be device memory"); }
    }
    if (!inventory) {
        const void* more[] = {prev, out->diag};
        for (const void* ptr : more) {
            if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly: buffers must be device memory"); }
        }
        if (g->max_degree > 0 && !is_device_ptr(out->offdiag)) {
            return fail(MPMC_E_INVALID_ARGUMENT, "assembly: offdiag must be device memory");
        }
    }
    need = 0;
    switch (nc) {
#define MPMC_CASE(N) case N: need = mpmc::dev::asm_record_bytes<N>(g->ncell); break;
        MPMC_FOR_EACH_NC2(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "assembly: unsupported nc");
    }
    return MPMC_OK;
}

// The kernels of one Newton evaluation with the given record workspace and error flag.
static hipError_t assembly_launch(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md,
                                  const double* q, const int32_t* branch, const double* prev, double dt,
                                  const mpmc_assembly_outputs* out, void* rec, unsigned long long* flag,
                                  bool inventory, hipStream_t s) {
    mpmc::dev::AsmModelDev dm{};
    dm.mu_a0 = md->mu_a0; dm.mu_a1 = md->mu_a1; dm.cp = md->cp; dm.t0 = md->t0; dm.cv_rock = md->cv_rock;
    for (int i = 0; i < MPMC_MAX_COMPONENTS; ++i) { dm.molar_mass[i] = md->molar_mass[i]; }
    const mpmc::dev::AsmGrid G{g->ncell, g->nface,    g->bulk,     g->poro,      g->owner,     g->neighbour,
                               g->trans, g->gdot,     g->gth,      g->adj_off,   g->adj_face,  g->slot_face,
                               g->slot_other, g->max_degree, g->block_stride};
    const mpmc::dev::AsmIo io{q, branch, prev, dt, out->cell_status, out->residual, out->diag, out->offdiag,
                              static_cast<double*>(rec), flag};
    hipError_t e = hipErrorInvalidValue;
    switch (m->nc) {
#define MPMC_CASE(N) case N: e = mpmc::dev::launch_assembly<N>(m->dev_constants, dm, G, io, inventory, s); break;
        MPMC_FOR_EACH_NC2(MPMC_CASE)
#undef MPMC_CASE
    default: break;
    }
    return e;
}

static int assembly_impl(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md, const double* q,
                         const int32_t* branch, const double* prev, double dt, const mpmc_assembly_outputs* out,
                         int32_t* domain_error, bool inventory, void* stream) {
    size_t need = 0;
    const int rc = assembly_check(m, g, md, q, branch, prev, dt, out, inventory, need);
    if (rc != MPMC_OK) { return rc; }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (g->rec_bytes < need) {
        if (g->rec) {
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) { return hip_fail(e, "assembly workspace sync"); }
            (void)hipFree(g->rec);
            g->rec = nullptr;
            g->rec_bytes = 0;
        }
        e = hipMalloc(&g->rec, need);
        if (e != hipSuccess) { g->rec = nullptr; return fail(MPMC_E_NO_MEMORY, "assembly: workspace allocation failed"); }
        g->rec_bytes = need;
    }
    e = assembly_launch(m, g, md, q, branch, prev, dt, out, g->rec, g->flag, inventory, s);
    if (e != hipSuccess) { return hip_fail(e, "assembly launch"); }
    if (domain_error) {
        unsigned long long flag = 0;
        e = hipMemcpyAsync(&flag, g->flag, sizeof(flag), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) { e = hipStreamSynchronize(s); }
        if (e != hipSuccess) { return hip_fail(e, "assembly flag read"); }
        *domain_error = flag ? 1 : 0;
    }
    return MPMC_OK;
}

int mpmc_assembly_2p(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md, const double* q,
                     const int32_t* branch, const double* prev, double dt, const mpmc_assembly_outputs* out,
                     int32_t* domain_error, void* stream) {
    return assembly_impl(m, g, md, q, branch, prev, dt, out, domain_error, false, stream);
}

int mpmc_inventory_2p(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md, const double* q,
                      const int32_t* branch, int32_t* cell_status, double* inventory, int32_t* domain_error,
                      void* stream) {
    const mpmc_assembly_outputs out{cell_status, inventory, nullptr, nullptr};
    return assembly_impl(m, g, md, q, branch, nullptr, 1.0, &out, domain_error, true, stream);
}

// ---------------------------------------------------------------- Schur-reduced system
int mpmc_assembly_reduce_2p(mpmc_tpfa_grid_t g, int nc, const int32_t* secondary, const mpmc_assembly_outputs* in,
                            const mpmc_reduced_outputs* out, void* stream) {
    if (!g || !in || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_reduce: null argument"); }
    if (nc < 2 || nc > MPMC_MAX_COMPONENTS) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_reduce: nc must be 2..12"); }
    const int Q = 2 * nc + 1;
    mpmc::dev::RedColumns cols{};
    bool is_sec[2 * MPMC_MAX_COMPONENTS + 1] = {};
    for (int k = 0; k < nc; ++k) {
        // default: x_0 of component 0 and the independent compositions of phase 1
        const int col = secondary ? secondary[k] : (k == 0 ? 3 : 3 + (nc - 1) + (k - 1));
        if (col < 0 || col >= Q || is_sec[col]) {
            return fail(MPMC_E_INVALID_ARGUMENT, "assembly_reduce: secondary columns must be nc distinct q columns");
        }
        is_sec[col] = true;
    }
    for (int col = 0, s = 0, p = 0; col < Q; ++col) {
        if (is_sec[col]) { cols.sec[s++] = col; }
        else { cols.prim[p++] = col; }
    }
    const void* ptrs[] = {in->residual, in->diag, in->offdiag, out->status, out->residual, out->diag, out->offdiag,
                          out->back};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_reduce: buffers must be device memory"); }
    }
    if (g->ncell == 0) { return MPMC_OK; }
    const mpmc::dev::RedIo io{g->ncell, g->max_degree, g->slot_other, in->residual, in->diag, in->offdiag,
                              out->back, out->residual, out->diag, out->offdiag, out->status};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((g->ncell + 63) / 64, 65536));
    switch (nc) {
#define MPMC_CASE(N)                                                                                              \
    case N:                                                                                                       \
        hipLaunchKernelGGL(mpmc::dev::k_red_cell<N>, dim3(blocks), dim3(64), 0, s, cols, io);                     \
        hipLaunchKernelGGL(mpmc::dev::k_red_rows<N>, dim3(blocks), dim3(64), 0, s, cols, io);                     \
        break;
        MPMC_FOR_EACH_NC2(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "assembly_reduce: unsupported nc");
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "assembly_reduce launch");
}

// ---------------------------------------------------------------- Newton-evaluation plan
struct mpmc_assembly_plan {
    void* rec = nullptr;                 // the plan's own record workspace
    unsigned long long* flag = nullptr;  // the plan's domain-error flag
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int mpmc_assembly_plan_destroy(mpmc_assembly_plan_t pl) {
    if (!pl) { return MPMC_OK; }
    const hipError_t e = hipDeviceSynchronize();
    if (pl->exec) { (void)hipGraphExecDestroy(pl->exec); }
    if (pl->graph) { (void)hipGraphDestroy(pl->graph); }
    if (pl->rec) { (void)hipFree(pl->rec); }
    if (pl->flag) { (void)hipFree(pl->flag); }
    delete pl;
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "assembly_plan_destroy");
}

int mpmc_assembly_plan_create(mpmc_mixture_t m, mpmc_tpfa_grid_t g, const mpmc_assembly_model* md, const double* q,
                              const int32_t* branch, const double* prev, double dt,
                              const mpmc_assembly_outputs* out, mpmc_assembly_plan_t* plan) {
    if (!plan) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_plan_create: null handle"); }
    *plan = nullptr;
    size_t need = 0;
    const int rc = assembly_check(m, g, md, q, branch, prev, dt, out, false, need);
    if (rc != MPMC_OK) { return rc; }
    std::unique_ptr<mpmc_assembly_plan> pl(new mpmc_assembly_plan());
    hipError_t e = hipMalloc(&pl->rec, need);
    if (e == hipSuccess) { e = hipMalloc(&pl->flag, sizeof(unsigned long long)); }
    if (e != hipSuccess) {
        mpmc_assembly_plan_destroy(pl.release());
        return fail(MPMC_E_NO_MEMORY, "assembly_plan_create: device allocation failed");
    }
    // one eager evaluation (on the null stream: the capture stream only records, as in
    // mpmc_flash_plan_create), then the same launch sequence captured once
    hipStream_t cs = nullptr;
    e = assembly_launch(m, g, md, q, branch, prev, dt, out, pl->rec, pl->flag, false, nullptr);
    if (e == hipSuccess) { e = hipStreamSynchronize(nullptr); }
    if (e == hipSuccess) { e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking); }
    if (e == hipSuccess) { e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal); }
    if (e == hipSuccess) {
        const hipError_t rec = assembly_launch(m, g, md, q, branch, prev, dt, out, pl->rec, pl->flag, false, cs);
        e = hipStreamEndCapture(cs, &pl->graph);
        if (rec != hipSuccess) { e = rec; }
    }
    if (e == hipSuccess) { e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0); }
    if (cs) { (void)hipStreamDestroy(cs); }
    if (e != hipSuccess) {
        mpmc_assembly_plan_destroy(pl.release());
        return hip_fail(e, "assembly_plan_create capture");
    }
    *plan = pl.release();
    return MPMC_OK;
}

int mpmc_assembly_plan_launch(mpmc_assembly_plan_t pl, void* stream) {
    if (!pl || !pl->exec) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_plan_launch: null plan"); }
    const hipError_t e = hipGraphLaunch(pl->exec, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "assembly_plan_launch");
}

int mpmc_assembly_plan_domain_error(mpmc_assembly_plan_t pl, void* stream, int32_t* domain_error) {
    if (!pl || !domain_error) { return fail(MPMC_E_INVALID_ARGUMENT, "assembly_plan_domain_error: null argument"); }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long flag = 0;
    hipError_t e = hipMemcpyAsync(&flag, pl->flag, sizeof(flag), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) { e = hipStreamSynchronize(s); }
    if (e != hipSuccess) { return hip_fail(e, "assembly_plan_domain_error"); }
    *domain_error = flag ? 1 : 0;
    return MPMC_OK;
}

} // extern "C"
