// Continuous-batching CCSD server (no reference counterpart: Aquarius runs one
// task DAG at a time, src/task/task.cxx:495-588).
//
// `slots` worker threads, each owning one long-lived context (its own HIP
// stream).  A job reserves its estimated device bytes against the server
// budget when admitted and releases them when it finishes; admission scans
// the queue in arrival order and starts the first job that fits (backfill),
// except that once the oldest queued job has waited longer than
// max_bypass_s no later job may overtake it.  Jobs run to convergence with
// their integrals, amplitudes and DIIS history resident on the device; their
// kernels overlap on the GPU through the per-slot streams (and, with grouping
// on, same-key GEMMs of different jobs run as one grouped launch on the
// batcher's stream; see aq_server_set_grouping).  The slot contexts
// autotune their GEMMs: the first job of each (o, v) size meets new GEMM shape
// buckets and times the kernel variants once; later jobs reuse the choices.  Nothing mutable is
// shared between jobs, so each job's arithmetic is independent of the mix.
#include "internal.hpp"

#include "../../include/aqamd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

namespace aqamd {

struct Job {
    aq_job_spec spec{};
    aq_job_result res{};
    double t_queued = 0.0;
    // caller-supplied integrals (spec.inputs / spec.fcidump point here after submit)
    std::vector<aq_tensor*> inputs;
    std::string fcidump;
};

// Shapes the solver expects of the ten inputs (aq_synth_build order).
inline std::vector<std::vector<int64_t>> input_shapes(int64_t o, int64_t v) {
    return {{v, o}, {v, v}, {o, o}, {v, v, v, v}, {v, v, v, o}, {v, v, o, o}, {v, o, v, o}, {v, o, o, v},
            {v, o, o, o}, {o, o, o, o}};
}

}  // namespace aqamd

using namespace aqamd;

size_t job_bytes(int o, int v, int nQ, int diis_order, int precision, bool integrals);

struct aq_server {
    int device = 0;
    size_t budget = 0, reserved = 0;
    double max_bypass_s = 0.0;
    std::vector<aq_ctx*> ctxs;
    GemmBatcher* batcher = nullptr;   // cross-job GEMM grouping (null: off)
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::map<int64_t, Job> jobs;
    int64_t next_id = 1;
    int64_t completed = 0, failed = 0, running = 0;
    bool stopping = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    // Pick the next admissible job (mu held); -1 if none.
    int64_t pick() {
        if (queue.empty()) return -1;
        const int64_t head = queue.front();
        const Job& hj = jobs[head];
        const bool head_blocked = now() - hj.t_queued > max_bypass_s;
        for (int64_t id : queue) {
            const Job& j = jobs[id];
            if (reserved + j.res.bytes_reserved <= budget) return id;
            if (head_blocked) return -1;   // the oldest job must go first
        }
        return -1;
    }

    // Runs one job on the slot's context; fills `job.res` fields of a private
    // copy (published under the lock by the caller).
    void run_job(int slot, Job& job) {
        aq_ctx* ctx = ctxs[slot];
        const aq_job_spec& s = job.spec;
        const int nQ = s.nQ > 0 ? s.nQ : 3 * (s.o + s.v);
        ctx->budget = job.res.bytes_reserved;
        ctx->peak_bytes = ctx_in_use(ctx);
        aq_tensor* ints[10] = {};        // owned by the job (synthetic or FCIDUMP)
        aq_tensor* const* use = ints;    // what the solver reads
        aq_rhfccsd* cc = nullptr;
        aq_tensor* mo = nullptr;
        auto cleanup = [&]() {
            if (cc) aq_rhfccsd_destroy(cc);
            if (mo) aq_tensor_destroy(mo);
            for (auto* t : ints)
                if (t) aq_tensor_destroy(t);
            aq_ctx_trim(ctx);
        };
        try {
            if (!job.inputs.empty()) {
                use = job.inputs.data();
            } else if (!job.fcidump.empty()) {
                // the reference's fcidump task: read the file, build the
                // closed-shell Fock matrix and the Dirac blocks on the device
                int norb = 0, nelec = 0, ms2 = 0;
                tensor::check(aq_fcidump_header(job.fcidump.c_str(), &norb, &nelec, &ms2));
                const int64_t n = norb;
                std::vector<double> h(size_t(n * n)), eri(size_t(n * n * n * n));
                double ecore = 0.0, escf = 0.0;
                tensor::check(aq_fcidump_read(job.fcidump.c_str(), 0, h.data(), eri.data(), &ecore));
                const int64_t len[4] = {n, n, n, n};
                tensor::check(aq_tensor_create(ctx, 4, len, 0, &mo));
                tensor::check(aq_write_all(ctx, mo, eri.data()));
                std::vector<double>().swap(eri);
                tensor::check(aq_mo_rhf_blocks(ctx, mo, nelec / 2, h.data(), &escf, nullptr, ints));
                aq_tensor_destroy(mo);
                mo = nullptr;
            } else {
                std::vector<double> eps(size_t(s.o + s.v)), B(size_t(nQ) * (s.o + s.v) * (s.o + s.v));
                tensor::check(aq_synth_host(s.seed, s.o, s.v, nQ, s.sigma, eps.data(), B.data()));
                tensor::check(aq_synth_build(ctx, s.o, s.v, nQ, eps.data(), B.data(), ints));
            }
            tensor::check(aq_rhfccsd_create(ctx, use, use + 3, &s.cfg, &cc));
            job.res.mp2 = 0.0;
            tensor::check(aq_rhfccsd_mp2(cc, &job.res.mp2));
            struct Count {
                int reduced = 0;
            } count;
            auto cb = [](int, double, double, double, int precision, void* user) {
                if (precision != AQ_PREC_FP64) static_cast<Count*>(user)->reduced += 1;
            };
            tensor::check(aq_rhfccsd_run(cc, cb, &count, &job.res.energy, &job.res.conv, &job.res.iterations,
                                         &job.res.converged));
            job.res.reduced_iterations = count.reduced;
            job.res.bytes_peak = ctx->peak_bytes;
        } catch (...) {
            job.res.bytes_peak = ctx->peak_bytes;
            cleanup();
            throw;
        }
        cleanup();
    }

    void worker(int slot) {
        (void)hipSetDevice(device);
        for (;;) {
            int64_t id = -1;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || (id = pick()) >= 0; });
                if (id < 0 && stopping) return;
                Job& j = jobs[id];
                queue.erase(std::find(queue.begin(), queue.end(), id));
                reserved += j.res.bytes_reserved;
                j.res.status = AQ_JOB_RUNNING;
                j.res.slot = slot;
                j.res.t_start = now();
                running += 1;
            }
            Job local;
            {
                std::lock_guard<std::mutex> lk(mu);
                local = jobs[id];
            }
            int code = AQ_OK;
            std::string msg;
            try {
                run_job(slot, local);
            } catch (const Error& e) {
                code = e.code;
                msg = e.what();
            } catch (const tensor::TensorError& e) {
                code = e.code;
                msg = e.what();
            } catch (const std::exception& e) {
                code = AQ_ERR_INVALID;
                msg = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                Job& j = jobs[id];
                j.res = local.res;
                j.res.t_end = now();
                reserved -= j.res.bytes_reserved;
                running -= 1;
                if (code == AQ_OK) {
                    j.res.status = AQ_JOB_DONE;
                    completed += 1;
                } else {
                    j.res.status = AQ_JOB_FAILED;
                    j.res.error_code = code;
                    std::strncpy(j.res.error, msg.c_str(), sizeof(j.res.error) - 1);
                    failed += 1;
                }
            }
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }
};

extern "C" {

size_t aq_job_bytes_estimate(int o, int v, int nQ, int diis_order, int precision) {
    return job_bytes(o, v, nQ, diis_order, precision, true);
}

}  // extern "C"

// Device bytes of one job.  integrals = false: the ten input tensors are the
// caller's (not counted).  nQ < 0: the integrals come from an FCIDUMP file of
// -nQ orbitals (the full MO tensor is on the device while the blocks are cut).
size_t job_bytes(int o, int v, int nQ, int diis_order, int precision, bool integrals) {
    if (o <= 0 || v <= 0) return 0;
    const bool from_file = nQ < 0;
    const double n4 = from_file ? std::pow(double(-nQ), 4) : 0.0;
    if (nQ <= 0) nQ = 3 * (o + v);
    const double O = o, V = v, Q = nQ, k = std::max(diis_order, 1);
    const double V4 = V * V * V * V, V3 = V * V * V * O, N2 = O * O * V * V, O3V = O * O * O * V, O4 = O * O * O * O;
    const double ints = V4 + V3 + 3 * N2 + O3V + O4 + V * V + O * O + V * O;
    // + the packed ladder: V+- (~V^4 / 2) and T+-, S+- (~N2 together)
    const double solver = 9 * N2 + 2 * O3V + O4 + 4 * V * V + 0.5 * V4 + N2;
    const double diis = 2 * k * (N2 + V * O);
    const double cache64 = 3 * V3 + 12 * N2;
    const double cacheR = precision != AQ_PREC_FP64 ? 0.5 * (V4 + 3 * V3 + 12 * N2) : 0.0;
    const double temps = 3 * N2;
    const double solve = (integrals ? ints : 0.0) + solver + diis + cache64 + cacheR + temps;
    const double build = !integrals ? 0.0 : from_file ? ints + n4 : ints + V4 + (V * V + V * O + O * O) * Q;
    return size_t(8.0 * std::max(solve, build) * 1.10) + (size_t(64) << 20);
}

extern "C" {

int aq_server_create(int device, size_t hbm_budget, int slots, double max_bypass_s, aq_server** out) {
    return guarded([&] {
        AQ_CHECK(out && slots >= 1 && slots <= 16, AQ_ERR_INVALID, "slots must be 1..16");
        auto* s = new aq_server();
        s->device = device;
        s->max_bypass_s = max_bypass_s;
        try {
            for (int i = 0; i < slots; ++i) {
                aq_ctx* c = nullptr;
                tensor::check(aq_ctx_create(device, 0, &c));
                s->ctxs.push_back(c);
                tensor::check(aq_ctx_set_autotune(c, 1));
            }
            size_t free_b = 0, total_b = 0;
            AQ_HIP(hipMemGetInfo(&free_b, &total_b));
            s->budget = hbm_budget ? std::min(hbm_budget, size_t(0.95 * double(free_b))) : size_t(0.9 * double(free_b));
            for (int i = 0; i < slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
        } catch (...) {
            for (auto* c : s->ctxs) aq_ctx_destroy(c);
            delete s;
            throw;
        }
        *out = s;
    });
}

int aq_server_submit(aq_server* s, const aq_job_spec* spec, int64_t* job_id) {
    return guarded([&] {
        AQ_CHECK(s && spec && job_id, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(!(spec->inputs && spec->fcidump), AQ_ERR_INVALID, "a job takes inputs or an FCIDUMP path, not both");
        aq_job_spec js = *spec;
        std::vector<aq_tensor*> inputs;
        std::string fcidump;
        size_t need = 0;
        if (spec->inputs) {
            for (int k = 0; k < 10; ++k) {
                AQ_CHECK(spec->inputs[k], AQ_ERR_INVALID, "job input tensor missing");
                AQ_CHECK(spec->inputs[k]->ctx && spec->inputs[k]->ctx->device == s->device, AQ_ERR_INVALID,
                         "job input tensors must live on the server's device");
                inputs.push_back(spec->inputs[k]);
            }
            AQ_CHECK(inputs[0]->ndim == 2, AQ_ERR_INVALID, "fAI must be [v,o]");
            js.o = int(inputs[0]->len[1]);
            js.v = int(inputs[0]->len[0]);
            const auto shapes = input_shapes(js.o, js.v);
            for (int k = 0; k < 10; ++k)
                AQ_CHECK(inputs[size_t(k)]->len == shapes[size_t(k)], AQ_ERR_INVALID,
                         "job input tensor " + std::to_string(k) + " has the wrong shape");
            AQ_CHECK(js.o > 0 && js.v > 0, AQ_ERR_INVALID, "job needs o > 0 and v > 0");
            // the integrals are borrowed: reserve the solver's own memory only
            need = job_bytes(js.o, js.v, 0, js.cfg.diis_order, js.cfg.precision, /*integrals=*/false);
            // their writes must be complete before another stream reads them
            for (auto* t : inputs) AQ_HIP(hipStreamSynchronize(t->ctx->stream));
        } else if (spec->fcidump) {
            fcidump = spec->fcidump;
            int norb = 0, nelec = 0, ms2 = 0;
            tensor::check(aq_fcidump_header(fcidump.c_str(), &norb, &nelec, &ms2));
            AQ_CHECK(ms2 == 0 && nelec % 2 == 0 && nelec > 0 && 2 * norb > nelec, AQ_ERR_UNSUPPORTED,
                     "FCIDUMP job needs a closed shell (MS2 = 0, even NELEC < 2 NORB)");
            js.o = nelec / 2;
            js.v = norb - js.o;
            need = job_bytes(js.o, js.v, -norb, js.cfg.diis_order, js.cfg.precision, true);
        } else {
            AQ_CHECK(spec->o > 0 && spec->v > 0, AQ_ERR_INVALID, "job needs o > 0 and v > 0");
            need = aq_job_bytes_estimate(spec->o, spec->v, spec->nQ, spec->cfg.diis_order, spec->cfg.precision);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        AQ_CHECK(!s->stopping, AQ_ERR_INVALID, "server is stopping");
        AQ_CHECK(need <= s->budget, AQ_ERR_OOM,
                 "job needs " + std::to_string(need) + " B, more than the server budget " + std::to_string(s->budget));
        const int64_t id = s->next_id++;
        Job& j = s->jobs[id];
        j.spec = js;
        j.inputs = std::move(inputs);
        j.fcidump = std::move(fcidump);
        j.spec.inputs = j.inputs.empty() ? nullptr : j.inputs.data();
        j.spec.fcidump = j.fcidump.empty() ? nullptr : j.fcidump.c_str();
        j.res.id = id;
        j.res.status = AQ_JOB_QUEUED;
        j.res.slot = -1;
        j.res.bytes_reserved = need;
        j.res.t_submit = s->now();
        j.t_queued = j.res.t_submit;
        s->queue.push_back(id);
        *job_id = id;
        s->cv_work.notify_all();
    });
}

int aq_server_poll(aq_server* s, int64_t id, aq_job_result* out) {
    return guarded([&] {
        AQ_CHECK(s && out, AQ_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        AQ_CHECK(it != s->jobs.end(), AQ_ERR_INVALID, "unknown job id");
        *out = it->second.res;
    });
}

int aq_server_wait(aq_server* s, int64_t id, aq_job_result* out) {
    return guarded([&] {
        AQ_CHECK(s, AQ_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        AQ_CHECK(it != s->jobs.end(), AQ_ERR_INVALID, "unknown job id");
        s->cv_done.wait(lk, [&] { return it->second.res.status >= AQ_JOB_DONE; });
        if (out) *out = it->second.res;
    });
}

int aq_server_drain(aq_server* s) {
    return guarded([&] {
        AQ_CHECK(s, AQ_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    });
}

int aq_server_stats_get(aq_server* s, aq_server_stats* out) {
    return guarded([&] {
        AQ_CHECK(s && out, AQ_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        out->submitted = s->next_id - 1;
        out->completed = s->completed;
        out->failed = s->failed;
        out->queued = int64_t(s->queue.size());
        out->running = s->running;
        out->budget = s->budget;
        out->reserved = s->reserved;
        out->elapsed_s = s->now();
    });
}

int aq_server_set_grouping(aq_server* s, int window_us) {
    return guarded([&] {
        AQ_CHECK(s, AQ_ERR_INVALID, "null server");
        AQ_CHECK(window_us >= 0, AQ_ERR_INVALID, "grouping window must be >= 0 us");
        std::lock_guard<std::mutex> lk(s->mu);
        AQ_CHECK(s->queue.empty() && s->running == 0, AQ_ERR_INVALID,
                 "grouping can change only while no job is queued or running");
        GemmBatcher* nb = window_us > 0 ? batcher_create(s->device, window_us) : nullptr;
        for (auto* c : s->ctxs) c->batcher = nb;
        batcher_destroy(s->batcher);
        s->batcher = nb;
    });
}

int aq_server_grouping_stats(aq_server* s, int64_t* launches, int64_t* gemms, int* max_group, int64_t* solo) {
    return guarded([&] {
        AQ_CHECK(s, AQ_ERR_INVALID, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->batcher) {
            batcher_stats(s->batcher, launches, gemms, max_group, solo);
        } else {
            if (launches) *launches = 0;
            if (gemms) *gemms = 0;
            if (max_group) *max_group = 0;
            if (solo) *solo = 0;
        }
    });
}

int aq_server_destroy(aq_server* s) {
    return guarded([&] {
        if (!s) return;
        aq_server_drain(s);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
        }
        s->cv_work.notify_all();
        for (auto& t : s->workers) t.join();
        for (auto* c : s->ctxs) aq_ctx_destroy(c);
        batcher_destroy(s->batcher);
        delete s;
    });
}

}  // extern "C"
