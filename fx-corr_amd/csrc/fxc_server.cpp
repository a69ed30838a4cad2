// Continuous-batching job server over the sector path.  One worker thread
// admits queued jobs by device footprint (plan tables + scratch, input
// window, FP64 spectra) against an HBM budget and a slot count, spreads the
// admitted jobs over a few HIP streams, and enqueues each job's chunks in
// order on its stream (copy-in, fused sector kernels, spectra += chunk).
// Jobs on different streams overlap on the device; a job's delay-model
// frame index and spectra stay resident between its chunks.  Results are
// copied back through pinned memory behind an event, so completion needs no
// stream synchronisation.  The sum over a job's chunks is taken in chunk
// order with deterministic kernels, so a job's result does not depend on
// what else runs.
#include "../../include/fxcorr.h"
#include "fxc_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);
extern "C" const int32_t *fxc_internal_plan_error_word(const fxc_plan *p);

using fxc::kNumProducts;
using fxc::launch_add_f64;
using fxc::launch_fill_zero;

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct ChunkIn {
    const uint8_t *r1, *r2;
    int64_t nf;
};

enum JobState { QUEUED, ACTIVE, DONE, FAILED };

struct Job {
    uint64_t id = 0;
    fxc_job_desc desc{};
    std::vector<double> lev[2];
    std::vector<double> table;
    std::string key;                 // plan-cache key
    std::deque<ChunkIn> chunks;
    bool closed = false;
    JobState state = QUEUED;
    std::string err;
    fxc_plan *plan = nullptr;
    int stream = -1;
    void *win[2] = {nullptr, nullptr};   // device input window (host input)
    size_t win_cap[2] = {0, 0};
    double *d_acc = nullptr, *d_chunk = nullptr;
    double *h_result = nullptr;          // pinned
    size_t half = 0;
    int64_t frames = 0;
    bool final_enqueued = false;
    hipEvent_t done_ev = nullptr;
    std::vector<double> result;
    clk::time_point t_submit, t_admit, t_done;
    uint64_t footprint = 0;
};

// Plans are interchangeable when every field of the description matches
// (built field by field: struct padding carries no meaning).
std::string plan_key(const fxc_job_desc &d, const std::vector<double> *lev) {
    std::string k;
    auto put = [&k](const void *p, size_t n) { k.append(static_cast<const char *>(p), n); };
    const fxc_plan_desc &p = d.plan;
    put(&p.fft_len, sizeof p.fft_len);
    for (int i = 0; i < 2; ++i) {
        put(&p.st[i].bits, sizeof p.st[i].bits);
        put(&p.st[i].lsb, sizeof p.st[i].lsb);
        put(p.st[i].shuffle_internal, sizeof p.st[i].shuffle_internal);
        put(lev[i].data(), lev[i].size() * sizeof(double));
        k.push_back('|');
    }
    put(&p.a1s, sizeof p.a1s);
    put(&p.a1e, sizeof p.a1e);
    put(&p.a2s, sizeof p.a2s);
    put(&p.out_grid, sizeof p.out_grid);
    put(&p.df_hz, sizeof p.df_hz);
    put(&p.products, sizeof p.products);
    put(&p.w1, sizeof p.w1);
    put(&p.w2, sizeof p.w2);
    put(&p.bit_out, sizeof p.bit_out);
    put(&p.output_lsb, sizeof p.output_lsb);
    put(&p.precision, sizeof p.precision);
    return k;
}

}  // namespace

struct fxc_server {
    fxc_server_cfg cfg{};
    std::vector<hipStream_t> streams;
    std::vector<int> stream_jobs;
    std::mutex mu;
    std::condition_variable cv;       // wakes the worker
    std::condition_variable done_cv;  // wakes waiters
    bool stop = false;
    std::thread worker;
    uint64_t next_id = 1;
    std::deque<uint64_t> pending;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::multimap<std::string, fxc_plan *> idle_plans;
    std::map<std::string, std::pair<uint64_t, fxc_plan_info>> key_info;   // plan bytes, info
    uint64_t idle_bytes = 0;
    uint64_t active_bytes = 0;
    int active = 0;
    // stats
    std::vector<double> latencies;
    int64_t failed = 0;
    int max_conc = 0;
    uint64_t peak_bytes = 0;

    uint64_t in_use() const { return active_bytes + idle_bytes; }
    void note_peak() { peak_bytes = std::max(peak_bytes, in_use()); }

    static uint64_t plan_bytes(fxc_plan *p) {
        uint64_t b = 0;
        (void)fxc_plan_device_bytes(p, &b);
        return b;
    }

    void evict_idle(uint64_t need) {
        while (!idle_plans.empty() && in_use() + need > cfg.hbm_budget_bytes) {
            auto it = idle_plans.begin();
            idle_bytes -= plan_bytes(it->second);
            fxc_plan_destroy(it->second);
            idle_plans.erase(it);
        }
    }
    // Evicts idle plans of other descriptions until `need` fits.
    void evict_idle_except(const std::string &key, uint64_t need) {
        for (auto it = idle_plans.begin(); it != idle_plans.end();) {
            const uint64_t keep = idle_plans.count(key) ? key_info[key].first : 0;
            if (in_use() - keep + need <= cfg.hbm_budget_bytes) break;
            if (it->first == key) {
                ++it;
                continue;
            }
            idle_bytes -= plan_bytes(it->second);
            fxc_plan_destroy(it->second);
            it = idle_plans.erase(it);
        }
    }

    void fail_job(Job &j, const std::string &msg) {
        j.state = FAILED;
        j.err = msg;
        j.t_done = clk::now();
        failed++;
        done_cv.notify_all();
    }

    void free_job_device(Job &j) {
        for (int i = 0; i < 2; ++i)
            if (j.win[i]) (void)hipFree(j.win[i]);
        if (j.d_acc) (void)hipFree(j.d_acc);
        if (j.d_chunk) (void)hipFree(j.d_chunk);
        if (j.h_result) (void)hipHostFree(j.h_result);
        if (j.done_ev) (void)hipEventDestroy(j.done_ev);
        j.win[0] = j.win[1] = nullptr;
        j.d_acc = j.d_chunk = j.h_result = nullptr;
        j.done_ev = nullptr;
    }

    // Device bytes of a job besides its plan: spectra (job + chunk) and, for
    // host input, the input window of its largest queued chunk.
    static uint64_t job_extra(const Job &j, const fxc_plan_info &info) {
        uint64_t win = 0;
        if (!j.desc.input_on_device)
            for (const ChunkIn &c : j.chunks)
                win = std::max<uint64_t>(win, (uint64_t)c.nf * (uint64_t)(info.bpf1 + info.bpf2));
        return 2ull * kNumProducts * (uint64_t)info.half * sizeof(double) + win;
    }

    // Returns true when the job left the queue (admitted or failed).  A plan
    // description's device size is learnt from its first creation; later
    // admissions are checked against the budget before anything is allocated.
    bool try_admit(Job &j) {
        auto known = key_info.find(j.key);
        if (known != key_info.end()) {
            const uint64_t need = known->second.first + job_extra(j, known->second.second);
            const bool cached = idle_plans.count(j.key) != 0;
            const uint64_t have = in_use() - (cached ? known->second.first : 0);
            if (have + need > cfg.hbm_budget_bytes) {
                evict_idle_except(j.key, need);
                const uint64_t have2 = in_use() - (idle_plans.count(j.key) ? known->second.first : 0);
                if (have2 + need > cfg.hbm_budget_bytes) {
                    if (active == 0) {
                        fail_job(j, "job footprint exceeds the server's HBM budget");
                        return true;
                    }
                    return false;   // wait for a running job to finish
                }
            }
        }
        fxc_plan *plan = nullptr;
        uint64_t pb = 0;
        auto it = idle_plans.find(j.key);
        if (it != idle_plans.end()) {
            plan = it->second;
            pb = plan_bytes(plan);
            idle_plans.erase(it);
            idle_bytes -= pb;
        } else {
            fxc_plan_desc d = j.desc.plan;
            d.st[0].levels = j.lev[0].data();
            d.st[1].levels = j.lev[1].data();
            if (fxc_plan_create(&d, &plan) != FXC_OK) {
                fail_job(j, std::string("plan: ") + fxc_last_error());
                return true;
            }
            pb = plan_bytes(plan);
        }
        fxc_plan_info info;
        (void)fxc_plan_get_info(plan, &info);
        key_info[j.key] = {pb, info};
        j.half = (size_t)info.half;
        const uint64_t acc_bytes = (uint64_t)kNumProducts * j.half * sizeof(double);
        const uint64_t need = pb + job_extra(j, info);
        evict_idle(need);
        if (in_use() + need > cfg.hbm_budget_bytes) {
            // first sight of this description and it does not fit now
            fxc_plan_destroy(plan);
            if (active == 0) {
                fail_job(j, "job footprint exceeds the server's HBM budget");
                return true;
            }
            return false;
        }
        if (hipMalloc(&j.d_acc, acc_bytes) != hipSuccess ||
            hipMalloc(&j.d_chunk, acc_bytes) != hipSuccess ||
            hipHostMalloc(&j.h_result, acc_bytes + sizeof(double), 0) != hipSuccess ||
            hipEventCreateWithFlags(&j.done_ev, hipEventDisableTiming) != hipSuccess) {
            free_job_device(j);
            fxc_plan_destroy(plan);
            fail_job(j, "device allocation failed at admission");
            return true;
        }
        int best = 0;
        for (int s = 1; s < (int)streams.size(); ++s)
            if (stream_jobs[s] < stream_jobs[best]) best = s;
        j.stream = best;
        stream_jobs[best]++;
        if (launch_fill_zero(j.d_acc, acc_bytes, streams[best]) != hipSuccess) {
            stream_jobs[best]--;
            free_job_device(j);
            fxc_plan_destroy(plan);
            fail_job(j, "zero fill failed");
            return true;
        }
        j.plan = plan;
        j.footprint = pb + 2 * acc_bytes;   // the input window is counted as it grows
        active_bytes += j.footprint;
        j.state = ACTIVE;
        j.t_admit = clk::now();
        active++;
        max_conc = std::max(max_conc, active);
        note_peak();
        return true;
    }

    bool dispatch(Job &j) {
        bool did = false;
        hipStream_t s = streams[j.stream];
        fxc_plan_info info;
        (void)fxc_plan_get_info(j.plan, &info);
        while (!j.chunks.empty()) {
            ChunkIn c = j.chunks.front();
            j.chunks.pop_front();
            did = true;
            const uint8_t *d1 = c.r1, *d2 = c.r2;
            if (!j.desc.input_on_device) {
                const size_t b[2] = {(size_t)c.nf * info.bpf1, (size_t)c.nf * info.bpf2};
                for (int i = 0; i < 2; ++i) {
                    if (b[i] > j.win_cap[i]) {   // grow: earlier chunks may still read it
                        (void)hipStreamSynchronize(s);
                        if (j.win[i]) (void)hipFree(j.win[i]);
                        active_bytes -= j.win_cap[i];
                        j.footprint -= j.win_cap[i];
                        j.win[i] = nullptr;
                        j.win_cap[i] = 0;
                        if (hipMalloc(&j.win[i], b[i]) != hipSuccess) {
                            fail_active(j, "input window allocation failed");
                            return true;
                        }
                        j.win_cap[i] = b[i];
                        active_bytes += b[i];
                        j.footprint += b[i];
                        note_peak();
                    }
                }
                if (hipMemcpyAsync(j.win[0], c.r1, b[0], hipMemcpyHostToDevice, s) != hipSuccess ||
                    hipMemcpyAsync(j.win[1], c.r2, b[1], hipMemcpyHostToDevice, s) != hipSuccess) {
                    fail_active(j, "input copy failed");
                    return true;
                }
                d1 = (const uint8_t *)j.win[0];
                d2 = (const uint8_t *)j.win[1];
            }
            fxc_delay_cfg dc = j.desc.delay;
            dc.geom_table = j.table.empty() ? nullptr : j.table.data();
            dc.geom_table_len = (int64_t)j.table.size();
            if (fxc_sector_run_device(j.plan, d1, d2, j.desc.first_frame + j.frames, c.nf, &dc,
                                      j.d_chunk, nullptr, s) != FXC_OK) {
                fail_active(j, std::string("sector: ") + fxc_last_error());
                return true;
            }
            if (launch_add_f64(j.d_acc, j.d_chunk, (size_t)kNumProducts * j.half, s) != hipSuccess) {
                fail_active(j, "accumulate failed");
                return true;
            }
            j.frames += c.nf;
        }
        if (j.closed && !j.final_enqueued) {
            const size_t bytes = (size_t)kNumProducts * j.half * sizeof(double);
            // the plan's device error word lands behind the spectra
            int32_t *err = reinterpret_cast<int32_t *>(j.h_result + (size_t)kNumProducts * j.half);
            *err = 0;
            if (hipMemcpyAsync(j.h_result, j.d_acc, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(err, fxc_internal_plan_error_word(j.plan), sizeof(int32_t),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipEventRecord(j.done_ev, s) != hipSuccess) {
                fail_active(j, "result copy failed");
                return true;
            }
            j.final_enqueued = true;
            did = true;
        }
        return did;
    }

    void retire(Job &j, bool to_cache) {
        if (j.plan) {
            if (to_cache) {   // scratch may have grown in use: re-learn its size
                const uint64_t pb = plan_bytes(j.plan);
                key_info[j.key].first = pb;
                idle_plans.emplace(j.key, j.plan);
                idle_bytes += pb;
            } else {
                fxc_plan_destroy(j.plan);
            }
            j.plan = nullptr;
        }
        free_job_device(j);
        active_bytes -= j.footprint;
        stream_jobs[j.stream]--;
        active--;
    }

    void fail_active(Job &j, const std::string &msg) {
        (void)hipStreamSynchronize(streams[j.stream]);
        retire(j, false);
        fail_job(j, msg);
    }

    bool complete(Job &j) {
        if (!j.final_enqueued) return false;
        hipError_t q = hipEventQuery(j.done_ev);
        if (q == hipErrorNotReady) return false;
        if (q != hipSuccess) {
            fail_active(j, std::string("device error: ") + hipGetErrorString(q));
            return true;
        }
        const int32_t err = *reinterpret_cast<const int32_t *>(j.h_result + (size_t)kNumProducts * j.half);
        if (err != 0) {   // a workgroup's bounded hand-over wait ran out: the spectra are not valid
            fail_active(j, "device error word set: a sector kernel's hand-over wait ran out");
            return true;
        }
        j.result.assign(j.h_result, j.h_result + (size_t)kNumProducts * j.half);
        retire(j, true);
        j.state = DONE;
        j.t_done = clk::now();
        latencies.push_back(ms_between(j.t_submit, j.t_done));
        done_cv.notify_all();
        return true;
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool work = false;
            for (auto &kv : jobs)
                if (kv.second->state == ACTIVE && complete(*kv.second)) work = true;
            while (!pending.empty() && active < cfg.max_active_jobs) {
                Job &j = *jobs[pending.front()];
                if (!try_admit(j)) break;
                pending.pop_front();
                work = true;
            }
            for (auto &kv : jobs)
                if (kv.second->state == ACTIVE && dispatch(*kv.second)) work = true;
            if (stop && pending.empty() && active == 0) break;
            if (!work) {
                // device work in flight: poll soon; otherwise sleep until woken
                if (active > 0) cv.wait_for(lk, std::chrono::microseconds(100));
                else cv.wait_for(lk, std::chrono::milliseconds(50));
            }
        }
    }
};

static fxc_status sfail(fxc_status s, const std::string &m) { return fxc_internal_fail(s, m.c_str()); }

extern "C" fxc_status fxc_server_create(const fxc_server_cfg *cfg, fxc_server **out) {
    if (!cfg || !out) return sfail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (cfg->max_active_jobs < 1 || cfg->n_streams < 1 || cfg->n_streams > 4 ||
        cfg->hbm_budget_bytes == 0)
        return sfail(FXC_ERR_INVALID_ARG, "server config: max_active_jobs >= 1, n_streams 1..4, budget > 0");
    int dc = 0;
    if (hipGetDeviceCount(&dc) != hipSuccess || dc < 1) return sfail(FXC_ERR_HIP, "no HIP device");
    auto *s = new fxc_server();
    s->cfg = *cfg;
    for (int i = 0; i < cfg->n_streams; ++i) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return sfail(FXC_ERR_HIP, "stream creation failed");
        }
        s->streams.push_back(st);
        s->stream_jobs.push_back(0);
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FXC_OK;
}

extern "C" fxc_status fxc_server_destroy(fxc_server *s) {
    if (!s) return FXC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto &kv : s->jobs)   // queued/active jobs without more input complete as they are
            kv.second->closed = true;
    }
    s->cv.notify_all();
    s->worker.join();
    for (auto &kv : s->idle_plans) fxc_plan_destroy(kv.second);
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    delete s;
    return FXC_OK;
}

extern "C" fxc_status fxc_job_submit(fxc_server *s, const fxc_job_desc *jd, uint64_t *id) {
    if (!s || !jd || !id) return sfail(FXC_ERR_INVALID_ARG, "null argument");
    auto j = std::make_unique<Job>();
    j->desc = *jd;
    for (int i = 0; i < 2; ++i) {
        const fxc_station_desc &st = jd->plan.st[i];
        if (!st.levels || st.nlevels <= 0) return sfail(FXC_ERR_INVALID_ARG, "station levels missing");
        j->lev[i].assign(st.levels, st.levels + st.nlevels);
        j->desc.plan.st[i].levels = nullptr;
    }
    if (jd->delay.geom_table && jd->delay.geom_table_len > 0)
        j->table.assign(jd->delay.geom_table, jd->delay.geom_table + jd->delay.geom_table_len);
    j->desc.delay.geom_table = nullptr;
    j->key = plan_key(j->desc, j->lev);
    j->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->id = s->next_id++;
        *id = j->id;
        s->pending.push_back(j->id);
        s->jobs[j->id] = std::move(j);
    }
    s->cv.notify_all();
    return FXC_OK;
}

extern "C" fxc_status fxc_job_push(fxc_server *s, uint64_t id, const uint8_t *r1, const uint8_t *r2,
                                   int64_t nf) {
    if (!s || !r1 || !r2 || nf <= 0) return sfail(FXC_ERR_INVALID_ARG, "bad push arguments");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return sfail(FXC_ERR_INVALID_ARG, "unknown job");
        Job &j = *it->second;
        if (j.closed) return sfail(FXC_ERR_INVALID_ARG, "job input already closed");
        if (j.state == FAILED) return sfail(FXC_ERR_INVALID_ARG, "job failed: " + j.err);
        j.chunks.push_back(ChunkIn{r1, r2, nf});
    }
    s->cv.notify_all();
    return FXC_OK;
}

extern "C" fxc_status fxc_job_close(fxc_server *s, uint64_t id) {
    if (!s) return sfail(FXC_ERR_INVALID_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return sfail(FXC_ERR_INVALID_ARG, "unknown job");
        it->second->closed = true;
    }
    s->cv.notify_all();
    return FXC_OK;
}

extern "C" fxc_status fxc_job_wait(fxc_server *s, uint64_t id, double timeout_s, int32_t *done) {
    if (!s || !done) return sfail(FXC_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(FXC_ERR_INVALID_ARG, "unknown job");
    Job *j = it->second.get();
    auto ready = [j] { return j->state == DONE || j->state == FAILED; };
    if (timeout_s < 0) s->done_cv.wait(lk, ready);
    else s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    *done = ready() ? 1 : 0;
    return FXC_OK;
}

extern "C" fxc_status fxc_job_result(fxc_server *s, uint64_t id, double *acc, fxc_job_stats *st) {
    if (!s) return sfail(FXC_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(FXC_ERR_INVALID_ARG, "unknown job");
    Job &j = *it->second;
    if (j.state == FAILED) return sfail(FXC_ERR_HIP, "job failed: " + j.err);
    if (j.state != DONE) return sfail(FXC_ERR_INVALID_ARG, "job not finished");
    if (acc) memcpy(acc, j.result.data(), j.result.size() * sizeof(double));
    if (st) {
        st->queue_ms = ms_between(j.t_submit, j.t_admit);
        st->latency_ms = ms_between(j.t_submit, j.t_done);
        st->frames = j.frames;
        st->footprint_bytes = j.footprint;
    }
    return FXC_OK;
}

extern "C" fxc_status fxc_job_release(fxc_server *s, uint64_t id) {
    if (!s) return sfail(FXC_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(FXC_ERR_INVALID_ARG, "unknown job");
    if (it->second->state != DONE && it->second->state != FAILED)
        return sfail(FXC_ERR_INVALID_ARG, "job still running");
    s->jobs.erase(it);
    return FXC_OK;
}

extern "C" fxc_status fxc_server_get_stats(fxc_server *s, fxc_server_stats *o) {
    if (!s || !o) return sfail(FXC_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto pct = [&v](double p) {   // nearest rank
        if (v.empty()) return 0.0;
        size_t r = (size_t)std::ceil(p / 100.0 * (double)v.size());
        return v[std::min(v.size(), std::max<size_t>(r, 1)) - 1];
    };
    o->jobs_done = (int64_t)v.size();
    o->jobs_failed = s->failed;
    o->p50_latency_ms = pct(50.0);
    o->p99_latency_ms = pct(99.0);
    double sum = 0.0;
    for (double x : v) sum += x;
    o->mean_latency_ms = v.empty() ? 0.0 : sum / (double)v.size();
    o->max_concurrent_jobs = s->max_conc;
    o->peak_hbm_bytes = s->peak_bytes;
    return FXC_OK;
}
