// Continuous-batching PT-flash server: the request loop of the reference's PT service
// (modules/runtime/include/mpmc/runtime/pt_service.hpp, one PtFlashBackend::solve per
// request, at most one concurrent solve by default, pt_grpc_adapter.hpp:23) rebuilt for
// one MI355X.
//
// Jobs (a mixture and a batch of cells in host memory) are admitted as they arrive. A
// launch ("chunk") packs cells of every pending job on one mixture: the mixture of the
// oldest waiting small job (at most `small_job_cells` cells) if there is one, else of the
// oldest large job; small jobs are packed first, then large jobs in arrival order, up to
// `small_job_cells` cells while small jobs wait and `chunk_cells` otherwise. Every chunk
// runs on one stream (the caller's): copy-in, the staged flash, copy-out. `slots` chunks
// are in flight at once, each with its own device chunk buffers and pinned host staging,
// so copies and the host gather/scatter of one chunk overlap the kernels of the other,
// and jobs that arrive meanwhile join the next chunk as soon as a slot frees. Device memory is fixed at creation (one flash workspace and the slot buffers),
// capped by `hbm_cap_bytes`; admission never allocates.
//
// One launch stream by design: the flash kernels that carry large private segments
// (the three-phase tail and the serial fallback, 8-22 KB per lane) were seen to abort
// with HSA_STATUS_ERROR_OUT_OF_RESOURCES when dispatched on additional hardware queues
// while the first queue held its scratch.
//
// Per-cell results are those of mpmc_flash_batch on the same cells: cells are
// independent, so chunking and interleaving do not change any result.
#include "../../include/mpmc_hnu_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <limits>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id;
    int nc;
    mpmc_mixture_t mixture;
    int64_t n;
    const double *p, *t, *z;
    int nhint = 0;                 // warm-start hints [nhint][nc][n] (host), nhint <= 2
    const double* hints = nullptr;
    mpmc_host_outputs out;
    Clock::time_point submitted;
    int64_t next_cell = 0;   // first cell not yet issued
    int64_t cells_done = 0;
    bool small = false;
    bool done = false;
    std::string error;       // first failure of a chunk holding cells of this job
    double latency_s = 0;
};

struct Slot {
    hipEvent_t done = nullptr;
    void* dev = nullptr;      // chunk inputs and outputs
    void* host = nullptr;     // pinned staging of the same layout
    size_t bytes = 0;
    // in flight: segments [c0, c0 + len) of jobs sharing one mixture, packed in order
    struct Segment {
        std::shared_ptr<Job> job;
        int64_t c0, len;
    };
    std::vector<Segment> segs;
    int64_t len = 0;
    bool busy() const { return !segs.empty(); }
};

// Chunk buffer layout for `len` cells of an nc-component mixture (doubles then ints).
struct ChunkLayout {
    size_t p, t, z, hint, beta, zf, x, lp, status, count, branch, total;
    ChunkLayout(int nc, int64_t len) {
        size_t o = 0;
        auto take = [&](size_t b) { const size_t at = o; o += (b + 255) / 256 * 256; return at; };
        p = take(8 * len);
        t = take(8 * len);
        z = take(8 * nc * len);
        hint = take(8 * 2 * nc * len);   // warm-start hints [2][nc][len], NaN rows where absent
        beta = take(8 * 3 * len);
        zf = take(8 * 3 * len);
        x = take(8 * 3 * nc * len);
        lp = take(8 * 3 * nc * len);
        status = take(4 * len);
        count = take(4 * len);
        branch = take(4 * 3 * len);
        total = o;
    }
};

} // namespace

struct mpmc_server {
    mpmc_server_options opt{};
    std::vector<Slot> slots;
    hipStream_t stream = nullptr; // the one launch stream
    void* workspace = nullptr;    // flash workspace shared by the chunks (stream ordered)
    size_t workspace_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> small_q, large_q;
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    std::map<std::string, mpmc_mixture_t> mixtures;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    size_t device_bytes = 0;
    std::atomic<int64_t> cells_served{0};

    ~mpmc_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) { worker.join(); }
        (void)hipStreamSynchronize(stream);
        for (auto& s : slots) {
            if (s.done) { (void)hipEventDestroy(s.done); }
            if (s.dev) { (void)hipFree(s.dev); }
            if (s.host) { (void)hipHostFree(s.host); }
        }
        if (workspace) { (void)hipFree(workspace); }
        for (auto& kv : mixtures) { (void)mpmc_mixture_destroy(kv.second); }
    }

    // Gather the chunk's segments into pinned staging and launch copy-in, flash, copy-out.
    bool launch(Slot& s, mpmc_mixture_t mixture, int nc) {
        const int64_t len = s.len;
        const ChunkLayout L(nc, len);
        char* h = static_cast<char*>(s.host);
        char* d = static_cast<char*>(s.dev);
        int64_t at = 0;
        bool hinted = false;
        for (const auto& g : s.segs) { hinted = hinted || g.job->nhint > 0; }
        const double nan = std::numeric_limits<double>::quiet_NaN();
        for (const auto& g : s.segs) {
            const auto& j = g.job;
            std::memcpy(h + L.p + 8 * at, j->p + g.c0, 8 * g.len);
            std::memcpy(h + L.t + 8 * at, j->t + g.c0, 8 * g.len);
            for (int i = 0; i < nc; ++i) {
                std::memcpy(h + L.z + 8 * (i * len + at), j->z + i * j->n + g.c0, 8 * g.len);
            }
            if (hinted) {
                for (int k = 0; k < 2; ++k) {
                    if (k < j->nhint) {
                        for (int i = 0; i < nc; ++i) {
                            std::memcpy(h + L.hint + 8 * ((k * nc + i) * len + at), j->hints + (k * nc + i) * j->n + g.c0,
                                        8 * g.len);
                        }
                    } else {
                        double* row = reinterpret_cast<double*>(h + L.hint) + (k * nc + 0) * len + at;
                        for (int64_t c = 0; c < g.len; ++c) { row[c] = nan; }
                    }
                }
            }
            at += g.len;
        }
        const size_t in_bytes = hinted ? L.beta : L.hint; // p, t, z (and hints) are the leading regions
        hipError_t e = hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) { fail_msg = std::string("copy in: ") + hipGetErrorString(e); return false; }
        mpmc_flash_options fo;
        mpmc_flash_default_options(&fo);
        fo.fugacity_tolerance = opt.fugacity_tolerance;
        fo.max_phases = opt.max_phases;
        fo.precision = opt.precision;
        const mpmc_flash_outputs out{reinterpret_cast<int32_t*>(d + L.status), reinterpret_cast<int32_t*>(d + L.count),
                                     reinterpret_cast<double*>(d + L.beta), reinterpret_cast<double*>(d + L.zf),
                                     reinterpret_cast<int32_t*>(d + L.branch), reinterpret_cast<double*>(d + L.x),
                                     reinterpret_cast<double*>(d + L.lp), nullptr};
        const int rc =
            hinted ? mpmc_flash_batch_hinted(mixture, len, reinterpret_cast<double*>(d + L.p),
                                             reinterpret_cast<double*>(d + L.t), reinterpret_cast<double*>(d + L.z), 2,
                                             reinterpret_cast<double*>(d + L.hint), &fo, &out, workspace,
                                             workspace_bytes, stream)
                   : mpmc_flash_batch_ws(mixture, len, reinterpret_cast<double*>(d + L.p),
                                         reinterpret_cast<double*>(d + L.t), reinterpret_cast<double*>(d + L.z), &fo,
                                         &out, workspace, workspace_bytes, stream);
        if (rc != MPMC_OK) {
            fail_msg = std::string("flash: ") + mpmc_last_error();
            return false;
        }
        e = hipMemcpyAsync(h + L.beta, d + L.beta, L.total - L.beta, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) { e = hipEventRecord(s.done, stream); }
        if (e != hipSuccess) { fail_msg = std::string("copy out: ") + hipGetErrorString(e); return false; }
        return true;
    }

    // Scatter a completed chunk from staging into the SoA host outputs of its jobs.
    void scatter(Slot& s) {
        const int nc = s.segs.front().job->nc;
        const int64_t len = s.len;
        const ChunkLayout L(nc, len);
        const char* h = static_cast<const char*>(s.host);
        int64_t at = 0;
        for (const auto& g : s.segs) {
            const auto& j = g.job;
            const int64_t n = j->n, c0 = g.c0, m = g.len;
            std::memcpy(j->out.status + c0, h + L.status + 4 * at, 4 * m);
            std::memcpy(j->out.phase_count + c0, h + L.count + 4 * at, 4 * m);
            for (int k = 0; k < 3; ++k) {
                std::memcpy(j->out.beta + k * n + c0, h + L.beta + 8 * (k * len + at), 8 * m);
                std::memcpy(j->out.zfactor + k * n + c0, h + L.zf + 8 * (k * len + at), 8 * m);
                std::memcpy(j->out.branch + k * n + c0, h + L.branch + 4 * (k * len + at), 4 * m);
                for (int i = 0; i < nc; ++i) {
                    const int64_t row = k * nc + i;
                    std::memcpy(j->out.x + row * n + c0, h + L.x + 8 * (row * len + at), 8 * m);
                    std::memcpy(j->out.ln_phi + row * n + c0, h + L.lp + 8 * (row * len + at), 8 * m);
                }
            }
            at += m;
        }
    }

    // Fill a free slot (caller holds mu): the mixture of the oldest pending small job, else
    // of the oldest large job; then cells of every pending job on that mixture, small jobs
    // first, up to small_chunk cells while small jobs wait and chunk_cells otherwise.
    mpmc_mixture_t fill(Slot& s, int& nc) {
        for (auto* q : {&small_q, &large_q}) {
            while (!q->empty() && q->front()->next_cell >= q->front()->n) { q->pop_front(); }
        }
        const std::shared_ptr<Job> head = !small_q.empty() ? small_q.front()
                                          : (!large_q.empty() ? large_q.front() : nullptr);
        if (!head) { return nullptr; }
        const mpmc_mixture_t mix = head->mixture;
        nc = head->nc;
        const int64_t budget = small_q.empty() ? opt.chunk_cells : std::min(opt.chunk_cells, small_chunk);
        s.segs.clear();
        s.len = 0;
        for (auto* q : {&small_q, &large_q}) {
            for (auto& j : *q) {
                if (s.len >= budget) { break; }
                if (j->mixture != mix || j->next_cell >= j->n) { continue; }
                const int64_t take = std::min<int64_t>(budget - s.len, j->n - j->next_cell);
                s.segs.push_back({j, j->next_cell, take});
                j->next_cell += take;
                s.len += take;
            }
        }
        return mix;
    }

    // A chunk has ended (caller holds mu). On failure its jobs carry the message and their
    // wait/poll report MPMC_E_DEVICE; other jobs are unaffected.
    void complete(Slot& s, const std::string& failure) {
        for (const auto& g : s.segs) {
            auto& j = g.job;
            j->cells_done += g.len;
            if (failure.empty()) { cells_served += g.len; }
            else if (j->error.empty()) { j->error = failure; }
            if (j->cells_done >= j->n && !j->done) {
                j->done = true;
                j->latency_s = std::chrono::duration<double>(Clock::now() - j->submitted).count();
            }
        }
        s.segs.clear();
        s.len = 0;
        cv_done.notify_all();
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            bool progressed = false;
            for (auto& s : slots) {
                if (!s.busy()) { continue; }
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipErrorNotReady) { continue; }
                std::string failure;
                if (q != hipSuccess) {
                    failure = std::string("server: device failure: ") + hipGetErrorString(q);
                } else {
                    lk.unlock();
                    scatter(s);
                    lk.lock();
                }
                complete(s, failure);
                progressed = true;
            }
            for (auto& s : slots) {
                if (s.busy()) { continue; }
                int nc = 0;
                const mpmc_mixture_t mix = fill(s, nc);
                if (!mix) { break; }
                lk.unlock();
                const bool ok = launch(s, mix, nc);
                if (!ok) {
                    // copies already queued on the stream read and write this slot's pinned
                    // staging: let them end before the slot is refilled
                    (void)hipStreamSynchronize(stream);
                }
                lk.lock();
                if (!ok) { complete(s, std::string("server: launch failed: ") + fail_msg); }
                progressed = true;
            }
            bool busy = false;
            for (auto& s : slots) { busy = busy || s.busy(); }
            if (stop && !busy && small_q.empty() && large_q.empty()) { break; }
            if (!progressed) {
                if (busy) {
                    cv_work.wait_for(lk, std::chrono::microseconds(50));
                } else {
                    cv_work.wait(lk, [&] { return stop || !small_q.empty() || !large_q.empty(); });
                }
            }
        }
    }

    int device = 0;
    int64_t small_chunk = 65536; // chunk size while small jobs wait
    std::string fail_msg; // last launch failure (scheduler thread only)
};

namespace {
thread_local std::string g_server_error;
int server_fail(int code, const std::string& what) {
    g_server_error = what;
    return code;
}
} // namespace

extern "C" {

void mpmc_server_default_options(mpmc_server_options* o) {
    if (!o) { return; }
    std::memset(o, 0, sizeof(*o));
    o->slots = 2;
    o->precision = 0;
    o->max_phases = 3;
    o->fugacity_tolerance = 1e-11;
    o->chunk_cells = 262144;
    o->small_job_cells = 65536;
    o->hbm_cap_bytes = size_t{200} << 30;
}

int mpmc_server_create(const mpmc_server_options* options, mpmc_server_t* out) {
    if (!out) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_create: null handle"); }
    *out = nullptr;
    mpmc_server_options o;
    mpmc_server_default_options(&o);
    if (options) { o = *options; }
    if (o.slots < 1 || o.slots > 32 || o.chunk_cells < 1 || o.chunk_cells > (int64_t{1} << 24) ||
        o.small_job_cells < 0 || (o.max_phases != 2 && o.max_phases != 3) || (o.precision < 0 || o.precision > 2) ||
        !(o.fugacity_tolerance > 0)) {
        return server_fail(MPMC_E_INVALID_ARGUMENT, "server_create: invalid options");
    }
    auto* s = new (std::nothrow) mpmc_server();
    if (!s) { return server_fail(MPMC_E_NO_MEMORY, "server_create: host allocation"); }
    s->opt = o;
    s->stream = static_cast<hipStream_t>(o.stream);
    s->small_chunk = std::max<int64_t>(1, o.small_job_cells);
    (void)hipGetDevice(&s->device);
    // one flash workspace (chunks run in stream order) and per-slot chunk buffers, sized
    // for the largest mixture
    const size_t ws = mpmc_flash_workspace_bytes(MPMC_MAX_COMPONENTS, o.chunk_cells);
    const size_t buf = ChunkLayout(MPMC_MAX_COMPONENTS, o.chunk_cells).total;
    if (ws + buf > o.hbm_cap_bytes) {
        delete s;
        return server_fail(MPMC_E_NO_MEMORY, "server_create: one chunk exceeds the device memory cap");
    }
    const int fit = static_cast<int>(std::min<size_t>(static_cast<size_t>(o.slots), (o.hbm_cap_bytes - ws) / buf));
    hipError_t e = hipMalloc(&s->workspace, ws);
    s->workspace_bytes = ws;
    s->device_bytes = ws;
    s->slots.resize(fit);
    for (auto& sl : s->slots) {
        if (e == hipSuccess) { e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming); }
        if (e == hipSuccess) { e = hipMalloc(&sl.dev, buf); }
        if (e == hipSuccess) { e = hipHostMalloc(&sl.host, buf, hipHostMallocDefault); }
        sl.bytes = buf;
        s->device_bytes += buf;
    }
    if (e != hipSuccess) {
        const std::string msg = std::string("server_create: ") + hipGetErrorString(e);
        delete s;
        return server_fail(MPMC_E_NO_MEMORY, msg);
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return MPMC_OK;
}

int mpmc_server_destroy(mpmc_server_t s) {
    delete s;
    return MPMC_OK;
}

int mpmc_server_slots(mpmc_server_t s) { return s ? static_cast<int>(s->slots.size()) : 0; }

int mpmc_server_submit(mpmc_server_t s, int nc, const double* tc, const double* pc, const double* omega,
                       const double* kij, int64_t ncell, const double* p, const double* t, const double* z,
                       const mpmc_host_outputs* out, int64_t* job_id) {
    return mpmc_server_submit_hinted(s, nc, tc, pc, omega, kij, ncell, p, t, z, 0, nullptr, out, job_id);
}

int mpmc_server_submit_hinted(mpmc_server_t s, int nc, const double* tc, const double* pc, const double* omega,
                              const double* kij, int64_t ncell, const double* p, const double* t, const double* z,
                              int nhint, const double* hints, const mpmc_host_outputs* out, int64_t* job_id) {
    if (!s || !out || !job_id || !p || !t || !z) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_submit: null argument"); }
    if (nhint < 0 || nhint > 2 || (nhint > 0 && !hints)) {
        return server_fail(MPMC_E_INVALID_ARGUMENT, "server_submit: nhint must be 0..2 with host hints");
    }
    if (nc < 1 || nc > MPMC_MAX_COMPONENTS || ncell < 1 || ncell > (int64_t{1} << 31)) {
        return server_fail(MPMC_E_INVALID_ARGUMENT, "server_submit: nc or ncell out of range");
    }
    if (!out->status || !out->phase_count || !out->beta || !out->zfactor || !out->branch || !out->x || !out->ln_phi) {
        return server_fail(MPMC_E_INVALID_ARGUMENT, "server_submit: null output buffer");
    }
    // resident mixture, keyed by its parameters
    std::string key(reinterpret_cast<const char*>(&nc), sizeof(nc));
    key.append(reinterpret_cast<const char*>(tc), 8 * nc);
    key.append(reinterpret_cast<const char*>(pc), 8 * nc);
    key.append(reinterpret_cast<const char*>(omega), 8 * nc);
    key.append(reinterpret_cast<const char*>(kij), 8 * nc * nc);
    mpmc_mixture_t mix = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->mixtures.find(key);
        if (it != s->mixtures.end()) { mix = it->second; }
    }
    if (!mix) {
        const int rc = mpmc_mixture_create(nc, tc, pc, omega, kij, &mix);
        if (rc != MPMC_OK) { return server_fail(rc, std::string("server_submit: ") + mpmc_last_error()); }
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->mixtures.find(key);
        if (it != s->mixtures.end()) {
            (void)mpmc_mixture_destroy(mix);
            mix = it->second;
        } else {
            s->mixtures.emplace(key, mix);
        }
    }
    auto j = std::make_shared<Job>();
    j->nc = nc;
    j->mixture = mix;
    j->n = ncell;
    j->p = p;
    j->t = t;
    j->z = z;
    j->nhint = nhint;
    j->hints = hints;
    j->out = *out;
    j->small = ncell <= s->opt.small_job_cells;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->id = s->next_id++;
        j->submitted = Clock::now();
        s->jobs.emplace(j->id, j);
        (j->small ? s->small_q : s->large_q).push_back(j);
    }
    s->cv_work.notify_all();
    *job_id = j->id;
    return MPMC_OK;
}

int mpmc_server_wait(mpmc_server_t s, int64_t job_id, double* latency_s) {
    if (!s) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_wait: null server"); }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_wait: unknown job"); }
    auto j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    if (latency_s) { *latency_s = j->latency_s; }
    s->jobs.erase(job_id);
    if (!j->error.empty()) { return server_fail(MPMC_E_DEVICE, j->error); }
    return MPMC_OK;
}

int mpmc_server_poll(mpmc_server_t s, int64_t job_id, int32_t* done, double* latency_s) {
    if (!s || !done) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_poll: null argument"); }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) { return server_fail(MPMC_E_INVALID_ARGUMENT, "server_poll: unknown job"); }
    *done = it->second->done ? 1 : 0;
    if (it->second->done && latency_s) { *latency_s = it->second->latency_s; }
    if (it->second->done && !it->second->error.empty()) { return server_fail(MPMC_E_DEVICE, it->second->error); }
    return MPMC_OK;
}

size_t mpmc_server_device_bytes(mpmc_server_t s) { return s ? s->device_bytes : 0; }

const char* mpmc_server_last_error(void) { return g_server_error.c_str(); }

} // extern "C"
