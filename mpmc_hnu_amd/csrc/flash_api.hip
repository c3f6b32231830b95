// C ABI of include/mpmc_hnu_amd.h: mixture residency, batched flash and batched
// phase evaluation. Argument checks mirror the reference's input contracts; numerical
// outcomes are per-cell statuses written by the kernels.
#include "closure_device.hpp"
#include "flash_pipeline.hpp"

#include "../../include/mpmc_hnu_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace mpmc::dev {
MPMC_FOR_EACH_NC(MPMC_DECLARE_LAUNCH)
MPMC_FOR_EACH_NC(MPMC_DECLARE_PIPELINE)
MPMC_FOR_EACH_NC2(MPMC_DECLARE_CLOSURE)
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

static int flash_batch_impl(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                            const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                            size_t workspace_bytes, bool own_workspace, void* stream) {
    if (!m || !out) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: null mixture or outputs"); }
    if (ncell < 0 || ncell > (int64_t{1} << 31)) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: bad ncell"); }
    mpmc_flash_options o;
    mpmc_flash_default_options(&o);
    if (options) { o = *options; }
    if (o.max_phases != 2 && o.max_phases != 3) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: max_phases must be 2 or 3");
    }
    if (!std::isfinite(o.fugacity_tolerance) || !(o.fugacity_tolerance > 0)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: fugacity tolerance must be finite and positive");
    }
    if (o.precision != 0 && o.precision != 1) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: precision must be 0 or 1"); }
    if (o.schedule != 0 && o.schedule != 1) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: schedule must be 0 or 1"); }
    if (o.schedule == 1 && o.precision != 0) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: the serial schedule runs FP64 only");
    }
    if (ncell == 0) { return MPMC_OK; }
    const void* ptrs[] = {p, t, z, out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                          out->ln_phi};
    for (const void* ptr : ptrs) {
        if (!is_device_ptr(ptr)) { return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: buffers must be device memory"); }
    }
    if (out->counters && !is_device_ptr(out->counters)) {
        return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: counters must be device memory");
    }
    const mpmc::dev::FlashOut fo{out->status, out->phase_count, out->beta, out->zfactor, out->branch, out->x,
                                 out->ln_phi, out->counters};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipErrorInvalidValue;
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
    if (own_workspace) {
        if (m->workspace_bytes < need) {
            if (m->workspace) {
                // the previous workspace may still be read by work queued on this stream
                e = hipStreamSynchronize(s);
                if (e != hipSuccess) { return hip_fail(e, "flash_batch workspace sync"); }
                (void)hipFree(m->workspace);
                m->workspace = nullptr;
                m->workspace_bytes = 0;
            }
            e = hipMalloc(&m->workspace, need);
            if (e != hipSuccess) {
                m->workspace = nullptr;
                return fail(MPMC_E_NO_MEMORY, std::string("flash_batch: workspace allocation of ") +
                                                  std::to_string(need) + " bytes failed");
            }
            m->workspace_bytes = need;
        }
        workspace = m->workspace;
        workspace_bytes = m->workspace_bytes;
    } else {
        if (!workspace || workspace_bytes < need || !is_device_ptr(workspace)) {
            return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch_ws: workspace must be device memory of "
                                                 "mpmc_flash_workspace_bytes(nc, ncell) bytes");
        }
    }
    switch (m->nc) {
#define MPMC_CASE(N)                                                                                             \
    case N:                                                                                                      \
        e = mpmc::dev::launch_pipeline<N>(m->dev_constants, workspace, workspace_bytes, ncell, p, t, z,          \
                                          o.max_phases, o.fugacity_tolerance, o.precision, fo, s);               \
        break;
        MPMC_FOR_EACH_NC(MPMC_CASE)
#undef MPMC_CASE
    default: return fail(MPMC_E_INVALID_ARGUMENT, "flash_batch: unsupported nc");
    }
    return e == hipSuccess ? MPMC_OK : hip_fail(e, "flash_batch launch");
}

int mpmc_flash_batch(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                     const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* stream) {
    return flash_batch_impl(m, ncell, p, t, z, options, out, nullptr, 0, true, stream);
}

int mpmc_flash_batch_ws(mpmc_mixture_t m, int64_t ncell, const double* p, const double* t, const double* z,
                        const mpmc_flash_options* options, const mpmc_flash_outputs* out, void* workspace,
                        size_t workspace_bytes, void* stream) {
    return flash_batch_impl(m, ncell, p, t, z, options, out, workspace, workspace_bytes, false, stream);
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

} // extern "C"
