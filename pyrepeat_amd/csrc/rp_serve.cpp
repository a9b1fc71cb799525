// Serving loop: concurrent charge fits on one device under an HBM budget.
//
// The reference runs one fit per process (repeat.py:2636-2789: read the cube, build
// ChargeCalculator, fit_charges, print).  Here requests are queued in arrival order and
// admitted to a slot when a slot is idle and the budget holds the request's device bound.
// A slot is a worker thread with its own context (HIP stream) and one grow-only fit object
// (rp_fit_reset), so admitted fits run concurrently on the device and a slot that has
// served a request of some size serves any smaller one without device allocation.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "rp_internal.h"

namespace rp {
namespace {

// A request with every array copied out of the caller's memory.
struct Job {
    uint64_t ticket = 0;
    rp_system sys{};
    rp_options opt{};
    std::vector<int32_t> Z;
    std::vector<double> xyz;
    std::vector<uint8_t> esp;
    double q_tot = 0.0;
    int32_t n_reduced = 0;
    std::vector<int32_t> red_ptr, red_src, expand_src;
    std::vector<double> red_mult, weights, diag_add, rhs_add;
    bool has_diag = false, has_rhs = false;
    uint64_t bound = 0;
    // results
    bool done = false;
    rp_result res{};
    std::vector<double> q;

    rp_constraints constraints() const {
        rp_constraints c{};
        c.q_tot = q_tot;
        c.n_reduced = n_reduced;
        c.red_ptr = red_ptr.data();
        c.red_src = red_src.data();
        c.red_mult = red_mult.data();
        c.weights = weights.data();
        c.diag_add = has_diag ? diag_add.data() : nullptr;
        c.rhs_add = has_rhs ? rhs_add.data() : nullptr;
        c.expand_src = expand_src.data();
        return c;
    }
};

struct Slot {
    int id = 0;
    rp_ctx *ctx = nullptr;
    rp_fit *fit = nullptr;
    bool busy = false;
    uint64_t held = 0;      // device bytes the slot's fit keeps between requests
    uint64_t reserved = 0;  // bytes charged against the budget (held, or held + bound while busy)
    std::thread th;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int validate_request(const rp_request *r, std::string &err) {
    const rp_system &s = r->system;
    const rp_constraints &c = r->constraints;
    if (s.n_atoms < 1 || !s.atomic_number || !s.atom_xyz || !s.esp) {
        err = "request: system needs atoms, positions and ESP";
        return RP_EINVAL;
    }
    if (s.n_grid[0] < 1 || s.n_grid[1] < 1 || s.n_grid[2] < 1) {
        err = "request: empty grid";
        return RP_EINVAL;
    }
    if (s.esp_dtype != RP_F64 && s.esp_dtype != RP_F32) {
        err = "request: bad esp_dtype";
        return RP_EINVAL;
    }
    if (c.n_reduced < 1 || c.n_reduced > s.n_atoms || !c.red_ptr || !c.red_src || !c.red_mult || !c.weights ||
        !c.expand_src) {
        err = "request: incomplete constraints";
        return RP_EINVAL;
    }
    return RP_OK;
}

}  // namespace

// Upper bound of the device bytes one fit of this request allocates: vdW filter, Ewald tables,
// the chunk buffers of its precision (and the factorized reciprocal tables of the split form), the matvec of --stats, Gram and solve, with the kept
// count bounded by the grid size.  The fit's grow-only workspaces are shared between phases,
// so each is bounded by the largest size any phase asks of it.
uint64_t fit_bytes_bound(const Setup &s, int n_atoms, int esp_dtype, int precision, int n_reduced, int n_src) {
    const int64_t n_total = (int64_t)s.nx * s.ny * s.nz;
    const uint64_t K = s.kcoefs.size();
    const uint64_t S = s.shifts.size() / 3;
    const uint64_t N = n_atoms, M = n_total;
    uint64_t ws[32] = {};
    auto at = [&](int i, uint64_t n) { ws[i] = std::max(ws[i], std::max<uint64_t>(n, 16)); };
    // largest power of two s with rows / s >= 1024 and tiles * s < 2048: the row-split loops stop earlier
    auto split_cap = [](int64_t rows, int tiles) {
        uint64_t cap = 1;
        while ((int64_t)(cap * 2) * 1024 <= rows && (uint64_t)tiles * cap * 2 < 2048) cap *= 2;
        return cap;
    };
    const ChunkPlan c64 = plan_chunks(M, n_atoms, (int)K, RP_PREC_FP64);
    const uint64_t NP = c64.NP;
    // fp64 basis chunks of the plain path's GEMM: its accumulate and its matvec (the factorized forms'
    // matvec after their accumulate recomputes their own basis into their chunk buffer, or streams the
    // basis the refined form keeps)
    if (precision == RP_PREC_FP64) {
        at(0, 8 * (uint64_t)c64.rows * c64.KP);
        at(1, 8 * (uint64_t)c64.KP * NP);
        at(2, 8 * (uint64_t)c64.rows * NP);
    }
    if (precision == RP_PREC_FP64) at(3, 8 * split_cap(c64.rows, c64.n_tiles) * c64.n_tiles * 128 * 128);
    at(22, 8 * N);
    at(23, 8 * M);
    uint64_t sep = 0;
    if (precision != RP_PREC_FP64) {
        const ChunkPlan c = plan_chunks(M, n_atoms, (int)K, precision);
        // one fp64 chunk buffer (the refined form's streamed refinement recomputes into it too)
        at(2, 8 * (uint64_t)c.NP * c.rows);
        at(5, 8 * (uint64_t)c.NP);
        if (precision == RP_PREC_FP64_FACT) {
            at(3, 8 * split_cap(c.rows, c.n_tiles) * c.n_tiles * 128 * 128);   // fp64 Gram partials
        } else {
            // the row-split partials and the second halves of the items that balance the last round
            // (at most 8 x 64 of them: half the CUs of an XCD, on devices of up to 1024 CUs)
            const uint64_t slots = split_cap(c.rows, c.n_tiles) * c.n_tiles;
            at(3, 8 * (slots + std::min<uint64_t>(slots, 8 * 64)) * 64 * 128);
            at(4, (uint64_t)c.slices * c.NP * c.rows);
            at(6, 8 * (uint64_t)c.NP);
            at(9, 8 * (uint64_t)c.n_tiles);
        }
        // factorized reciprocal sum: k box, T1 per plane (7 x 13 complex per plane and atom), the
        // coefficients w e^{-i k.R} per atom and k-vector,
        // line table of the kept points, per-chunk line ranges
        const uint64_t nb = (uint64_t)ceil_div(M, 256), n_chunks = (uint64_t)ceil_div(M, c.rows);
        sep += 4 * 8 * 13 * 13 + (K > 0 ? 16 * 7 * 13 * (uint64_t)s.nz * N + 16 * (uint64_t)N * K : 0);
        sep += 4 * nb + 8 * (nb + 1) + 4 * (M + 1) + 4 * M + 8 * n_chunks + 16 * 7;
        if (precision == RP_PREC_OZAKI_REFINED) {
            // the streamed refinement: the residual e over the kept points and its block sums, the column
            // pass's partial sums.  Keeping the basis instead (ws[7]) is not in the bound: a fit takes it
            // only when its context's budget grows by that much (the serving loop asks its own budget).
            at(8, 8 * (M + 2 + (uint64_t)ceil_div(M, 512)));
            at(15, 8 * 4 * 256 * N);
        }
    }
    // bordered solve
    const uint64_t nr = n_reduced, n = nr + 1, ld = (uint64_t)round_up((int64_t)n + 1, 4);
    const uint64_t nsrc = std::max(n_src, 1);
    at(12, 8 * n * ld);
    // least-squares fallback of an exactly singular system: eigenvectors, round pairs and vectors
    at(11, 8 * n * n);
    at(14, (uint64_t)round_up((int64_t)(48 * ((n + 1) / 2)), 256) + 8 * (2 * n + 4) + 256);
    // packed constraint arrays: ints (red_ptr, red_src, expand_src, inv_ptr, inv_set), then doubles
    // (red_mult, weights, diag_add, rhs_add, inv_mult)
    at(13, (uint64_t)round_up((int64_t)(round_up((int64_t)(4 * ((nr + 1) + 2 * nsrc + 2 * N + 1)), 8) +
                                        8 * (2 * nsrc + 3 * nr)), 16));
    // refinement vectors and state
    at(10, 8 * (uint64_t)round_up((int64_t)(2 * N + 3 * nr + 2), 2) + 16);
    at(19, 4 * n);
    at(20, 8 * n * ld);
    for (int i = 21; i <= 23; ++i) at(i, 8 * nr);
    at(24, 4);
    at(25, 8 * n);
    at(27, 8 * N);
    at(28, 8);
    at(31, 4 * (1 + (uint64_t)ceil_div((int64_t)nr, 64)));
    at(29, 8 * nr);
    at(30, 8 * 64 * 64 * (uint64_t)ceil_div((int64_t)nr, 64));
    uint64_t b = filter_bytes_bound(s, esp_dtype) + sep;
    for (uint64_t w : ws) b += w;
    // device copy of the Ewald setup, atoms, augmented Gram
    b += 16 * (8 + 15 + 15) + 12 * K + 8 * K + 24 * K + 24 * N + 24 * S + 8 * (uint64_t)c64.n_tiles + 16 * 8;
    b += 24 * N + 8 * NP * NP;
    return b;
}

}  // namespace rp

using namespace rp;

struct rp_server {
    int device = 0;
    uint64_t budget = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;  // queue or reservations changed
    std::condition_variable cv_done;  // a request finished
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    std::vector<std::unique_ptr<Slot>> slots;
    bool stopping = false;
    uint64_t next_ticket = 1;
    int64_t submitted = 0, completed = 0, failed = 0;
    int running = 0, max_running = 0;
    uint64_t peak_reserved = 0;
    std::string err;

    uint64_t reserved_total() const {
        uint64_t t = 0;
        for (auto &s : slots) t += s->reserved;
        return t;
    }
    uint64_t busy_reserved() const {
        uint64_t t = 0;
        for (auto &s : slots)
            if (s->busy) t += s->reserved;
        return t;
    }

    void worker(Slot *slot);
    void run_job(Slot *slot, Job &job, bool trim_first);
};

void rp_server::run_job(Slot *slot, Job &job, bool trim_first) {
    rp_ctx *ctx = slot->ctx;
    // the slot's context may hold what admission reserved for it; a refined fit asks for more to keep
    // its basis, granted while the server's budget holds it
    {
        std::lock_guard<std::mutex> lk(mu);
        ctx->budget = slot->reserved;
    }
    ctx->grow_budget = [this, slot, ctx](uint64_t extra) {
        std::lock_guard<std::mutex> lk(mu);
        if (reserved_total() + extra > budget) return false;
        slot->reserved += extra;
        ctx->budget += extra;
        peak_reserved = std::max(peak_reserved, reserved_total());
        return true;
    };
    if (trim_first && slot->fit) {
        rp_fit_destroy(slot->fit);
        slot->fit = nullptr;
    }
    int rc;
    if (!slot->fit) {
        rc = rp_fit_create(ctx, &job.sys, &job.opt, &slot->fit);
    } else {
        rc = rp_fit_reset(slot->fit, &job.sys, &job.opt);
    }
    if (rc == RP_OK) rc = rp_fit_accumulate(slot->fit);
    job.q.assign(job.sys.n_atoms, 0.0);
    if (rc == RP_OK) {
        const rp_constraints con = job.constraints();
        int32_t fb = 0;
        rc = rp_fit_solve(slot->fit, &con, job.q.data(), nullptr, &fb);
    }
    job.res.status = rc;
    job.res.slot = slot->id;
    job.res.n_atoms = job.sys.n_atoms;
    if (slot->fit) {
        job.res.n_keep = slot->fit->n_keep;
        job.res.device_ms[0] = slot->fit->t_filter_ms;
        job.res.device_ms[1] = slot->fit->t_accum_ms;
        job.res.device_ms[2] = slot->fit->t_solve_ms;
    }
    if (rc != RP_OK) {
        std::snprintf(job.res.error, sizeof(job.res.error), "%s", ctx->err.c_str());
        // a failed fit may hold partial state: start the next request on this slot from scratch
        if (slot->fit) {
            rp_fit_destroy(slot->fit);
            slot->fit = nullptr;
        }
    }
}

void rp_server::worker(Slot *slot) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        std::shared_ptr<Job> job;
        bool trim_first = false;
        for (;;) {
            if (!queue.empty()) {
                const uint64_t need = queue.front()->bound;
                const uint64_t others = reserved_total() - slot->reserved;
                if (others + slot->held + need <= budget) {
                    job = queue.front();
                    slot->reserved = slot->held + need;  // buffers only grow: held + bound covers the run
                    break;
                }
                if (others + need <= budget) {
                    job = queue.front();
                    trim_first = true;
                    slot->reserved = need;
                    break;
                }
                // blocked: idle slots give their memory back when that alone lets the head in
                if (slot->held > 0 && busy_reserved() + need <= budget) {
                    slot->busy = true;  // keeps other workers from counting this slot as idle meanwhile
                    lk.unlock();
                    rp_fit_destroy(slot->fit);
                    slot->fit = nullptr;
                    lk.lock();
                    slot->busy = false;
                    slot->held = 0;
                    slot->reserved = 0;
                    cv_work.notify_all();
                    continue;
                }
            } else if (stopping) {
                return;
            }
            cv_work.wait(lk);
        }
        queue.pop_front();
        slot->busy = true;
        ++running;
        max_running = std::max(max_running, running);
        peak_reserved = std::max(peak_reserved, reserved_total());
        job->res.t_admit = seconds_since(t0);
        job->res.bytes_bound = job->bound;
        lk.unlock();

        run_job(slot, *job, trim_first);

        lk.lock();
        job->res.t_done = seconds_since(t0);
        uint64_t in_use = 0;
        rp_ctx_hbm_in_use(slot->ctx, &in_use);
        slot->held = in_use;
        slot->reserved = in_use;
        slot->busy = false;
        --running;
        ++completed;
        if (job->res.status != RP_OK) ++failed;
        job->done = true;
        // release the request's input copies; the result stays until polled
        std::vector<uint8_t>().swap(job->esp);
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int rp_server_create(int device, int n_slots, uint64_t hbm_budget_bytes, rp_server **out) {
    if (!out || n_slots < 1 || n_slots > 16 || hbm_budget_bytes == 0) {
        set_global_error("rp_server_create: need out, 1 <= n_slots <= 16 and a nonzero budget");
        return RP_EINVAL;
    }
    *out = nullptr;
    auto *srv = new rp_server();
    srv->device = device;
    srv->budget = hbm_budget_bytes;
    srv->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n_slots; ++i) {
        auto slot = std::make_unique<Slot>();
        slot->id = i;
        // each slot's context carries the whole budget as a hard cap; admission keeps the sum within it
        const int rc = rp_ctx_create(device, hbm_budget_bytes, &slot->ctx);
        if (rc != RP_OK) {
            for (auto &s : srv->slots) rp_ctx_destroy(s->ctx);
            delete srv;
            return rc;
        }
        srv->slots.push_back(std::move(slot));
    }
    for (auto &s : srv->slots) {
        Slot *sp = s.get();
        s->th = std::thread([srv, sp] { srv->worker(sp); });
    }
    *out = srv;
    return RP_OK;
}

int rp_server_destroy(rp_server *srv) {
    if (!srv) return RP_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    for (auto &s : srv->slots)
        if (s->th.joinable()) s->th.join();
    for (auto &s : srv->slots) {
        if (s->fit) rp_fit_destroy(s->fit);
        rp_ctx_destroy(s->ctx);
    }
    delete srv;
    return RP_OK;
}

const char *rp_server_error(const rp_server *srv) { return srv ? srv->err.c_str() : rp_last_error(nullptr); }

double rp_server_now(const rp_server *srv) { return srv ? seconds_since(srv->t0) : 0.0; }

int rp_request_bytes(const rp_request *req, uint64_t *bytes) {
    if (!req || !bytes) return RP_EINVAL;
    std::string err;
    int rc = validate_request(req, err);
    if (rc != RP_OK) {
        set_global_error(err);
        return rc;
    }
    Setup s;
    rc = build_setup(&req->system, &req->options, s, err);
    if (rc != RP_OK) {
        set_global_error(err);
        return rc;
    }
    const rp_constraints &c = req->constraints;
    *bytes = fit_bytes_bound(s, req->system.n_atoms, req->system.esp_dtype, req->options.precision, c.n_reduced,
                             c.red_ptr[c.n_reduced]);
    return RP_OK;
}

int rp_submit(rp_server *srv, const rp_request *req, uint64_t *ticket) {
    if (!srv || !req || !ticket) return RP_EINVAL;
    const double t_submit = seconds_since(srv->t0);
    uint64_t bound = 0;
    int rc = rp_request_bytes(req, &bound);
    if (rc != RP_OK) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->err = rp_last_error(nullptr);
        return rc;
    }
    if (bound > srv->budget) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->err = "request needs up to " + std::to_string(bound) + " device bytes, over the server budget of " +
                   std::to_string(srv->budget);
        return RP_EBUDGET;
    }
    auto job = std::make_shared<Job>();
    const rp_system &s = req->system;
    const rp_constraints &c = req->constraints;
    const int64_t n_total = (int64_t)s.n_grid[0] * s.n_grid[1] * s.n_grid[2];
    job->Z.assign(s.atomic_number, s.atomic_number + s.n_atoms);
    job->xyz.assign(s.atom_xyz, s.atom_xyz + 3 * (size_t)s.n_atoms);
    const size_t esp_bytes = (size_t)n_total * (s.esp_dtype == RP_F32 ? 4 : 8);
    job->esp.assign(static_cast<const uint8_t *>(s.esp), static_cast<const uint8_t *>(s.esp) + esp_bytes);
    job->sys = s;
    job->sys.atomic_number = job->Z.data();
    job->sys.atom_xyz = job->xyz.data();
    job->sys.esp = job->esp.data();
    job->opt = req->options;
    job->q_tot = c.q_tot;
    job->n_reduced = c.n_reduced;
    const int nsrc = c.red_ptr[c.n_reduced];
    job->red_ptr.assign(c.red_ptr, c.red_ptr + c.n_reduced + 1);
    job->red_src.assign(c.red_src, c.red_src + nsrc);
    job->red_mult.assign(c.red_mult, c.red_mult + nsrc);
    job->weights.assign(c.weights, c.weights + c.n_reduced);
    job->has_diag = c.diag_add != nullptr;
    job->has_rhs = c.rhs_add != nullptr;
    if (job->has_diag) job->diag_add.assign(c.diag_add, c.diag_add + c.n_reduced);
    if (job->has_rhs) job->rhs_add.assign(c.rhs_add, c.rhs_add + c.n_reduced);
    job->expand_src.assign(c.expand_src, c.expand_src + s.n_atoms);
    job->bound = bound;
    job->res.slot = -1;
    job->res.t_submit = t_submit;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stopping) return RP_ESTATE;
        job->ticket = srv->next_ticket++;
        srv->jobs[job->ticket] = job;
        srv->queue.push_back(job);
        ++srv->submitted;
        *ticket = job->ticket;
    }
    srv->cv_work.notify_all();
    return RP_OK;
}

int rp_poll(rp_server *srv, uint64_t ticket, int32_t wait_ms, rp_result *res, double *q_out) {
    if (!srv) return RP_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) return RP_EINVAL;
    std::shared_ptr<Job> job = it->second;
    if (!job->done) {
        if (wait_ms < 0)
            srv->cv_done.wait(lk, [&] { return job->done; });
        else if (wait_ms > 0)
            srv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return job->done; });
        if (!job->done) return RP_EAGAIN;
    }
    if (res) *res = job->res;
    if (q_out) std::copy(job->q.begin(), job->q.end(), q_out);
    srv->jobs.erase(ticket);
    return RP_OK;
}

int rp_server_stats_get(const rp_server *srv_c, rp_server_stats *out) {
    if (!srv_c || !out) return RP_EINVAL;
    rp_server *srv = const_cast<rp_server *>(srv_c);
    std::lock_guard<std::mutex> lk(srv->mu);
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->queued = (int32_t)srv->queue.size();
    out->running = srv->running;
    out->max_running = srv->max_running;
    out->n_slots = (int32_t)srv->slots.size();
    out->budget = srv->budget;
    out->reserved = srv->reserved_total();
    out->peak_reserved = srv->peak_reserved;
    return RP_OK;
}

}  // extern "C"
