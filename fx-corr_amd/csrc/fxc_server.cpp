// Continuous-batching job server over the sector path.  One worker thread
// admits queued jobs by device footprint (plan tables + scratch, input
// window, FP64 spectra) against an HBM budget and a slot count, spreads the
// admitted jobs over a few HIP streams, and enqueues each job's chunks in
// order on its stream (fused sector kernels, spectra += chunk), with host
// input copied in on a companion copy stream into two alternating windows.
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
    // host input: two device windows per station, used by alternate chunks,
    // so a chunk's copy (on the copy stream) overlaps the previous chunk's
    // kernels (on the job's stream); copy_ev[b] / comp_ev[b] order a window's
    // copy before its kernels and its kernels before the next copy into it
    void *win[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [window][station]
    size_t win_cap[2][2] = {{0, 0}, {0, 0}};
    hipEvent_t copy_ev[2] = {nullptr, nullptr}, comp_ev[2] = {nullptr, nullptr};
    bool win_used[2] = {false, false};
    int64_t nchunk = 0;
    double *d_acc = nullptr, *d_chunk = nullptr;
    size_t acc_cap = 0, chunk_cap = 0;   // their capacities (pooled buffers may be larger)
    double *h_result = nullptr;          // pinned
    size_t res_cap = 0;
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
    std::vector<hipStream_t> copy_streams;   // host-input copies, one per stream
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
    // Buffers and events of finished jobs, kept for later ones: releasing
    // device or pinned memory waits for the whole device, and creating
    // events costs a runtime call per job.  Device buffers count against
    // the HBM budget until they are released to make room.
    std::multimap<size_t, void *> dev_pool;
    std::multimap<size_t, double *> host_pool;
    std::vector<hipEvent_t> ev_pool;
    uint64_t pool_bytes = 0;

    uint64_t in_use() const { return active_bytes + idle_bytes + pool_bytes; }
    void note_peak() { peak_bytes = std::max(peak_bytes, in_use()); }

    // A device buffer of at least n bytes (a pooled one of at most 2n, else
    // a new one); *cap receives its size.
    void *dev_get(size_t n, size_t *cap) {
        auto it = dev_pool.lower_bound(n);
        if (it != dev_pool.end() && it->first <= 2 * n) {
            void *b = it->second;
            *cap = it->first;
            pool_bytes -= it->first;
            dev_pool.erase(it);
            return b;
        }
        void *b = nullptr;
        if (hipMalloc(&b, n) != hipSuccess) {
            trim_pool(~0ull);   // the pool may hold the memory
            if (hipMalloc(&b, n) != hipSuccess) return nullptr;
        }
        *cap = n;
        return b;
    }
    void dev_put(void *b, size_t cap) {
        if (!b) return;
        dev_pool.emplace(cap, b);
        pool_bytes += cap;
    }
    // Releases pooled device buffers (largest first) until `need` more
    // bytes fit the budget (all of them for need = ~0).
    void trim_pool(uint64_t need) {
        while (!dev_pool.empty() && (need == ~0ull || in_use() + need > cfg.hbm_budget_bytes)) {
            auto it = std::prev(dev_pool.end());
            (void)hipFree(it->second);
            pool_bytes -= it->first;
            dev_pool.erase(it);
        }
    }
    double *host_get(size_t n, size_t *cap) {
        auto it = host_pool.lower_bound(n);
        if (it != host_pool.end() && it->first <= 2 * n) {
            double *b = it->second;
            *cap = it->first;
            host_pool.erase(it);
            return b;
        }
        double *b = nullptr;
        if (hipHostMalloc(reinterpret_cast<void **>(&b), n, 0) != hipSuccess) return nullptr;
        *cap = n;
        return b;
    }
    hipEvent_t ev_get() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
    void release_pools() {
        trim_pool(~0ull);
        for (auto &kv : host_pool) (void)hipHostFree(kv.second);
        host_pool.clear();
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        ev_pool.clear();
    }

    static uint64_t plan_bytes(fxc_plan *p) {
        uint64_t b = 0;
        (void)fxc_plan_device_bytes(p, &b);
        return b;
    }

    void evict_idle(uint64_t need) {
        trim_pool(need);
        while (!idle_plans.empty() && in_use() + need > cfg.hbm_budget_bytes) {
            auto it = idle_plans.begin();
            idle_bytes -= plan_bytes(it->second);
            fxc_plan_destroy(it->second);
            idle_plans.erase(it);
        }
    }
    // Evicts idle plans of other descriptions until `need` fits.
    void evict_idle_except(const std::string &key, uint64_t need) {
        trim_pool(need);
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

    // Returns a job's buffers and event to the pools (its device work has
    // finished: complete() saw its event, or the failure path synchronised
    // its stream).
    void free_job_device(Job &j) {
        for (int b = 0; b < 2; ++b) {
            for (int i = 0; i < 2; ++i) {
                dev_put(j.win[b][i], j.win_cap[b][i]);
                j.win[b][i] = nullptr;
                j.win_cap[b][i] = 0;
            }
            if (j.copy_ev[b]) ev_pool.push_back(j.copy_ev[b]);
            if (j.comp_ev[b]) ev_pool.push_back(j.comp_ev[b]);
            j.copy_ev[b] = j.comp_ev[b] = nullptr;
            j.win_used[b] = false;
        }
        dev_put(j.d_acc, j.acc_cap);
        dev_put(j.d_chunk, j.chunk_cap);
        if (j.h_result) host_pool.emplace(j.res_cap, j.h_result);
        if (j.done_ev) ev_pool.push_back(j.done_ev);
        j.d_acc = j.d_chunk = j.h_result = nullptr;
        j.acc_cap = j.chunk_cap = j.res_cap = 0;
        j.done_ev = nullptr;
    }

    // Device bytes of a job besides its plan: spectra (job + chunk) and, for
    // host input, the two input windows of its largest queued chunk.
    static uint64_t job_extra(const Job &j, const fxc_plan_info &info) {
        uint64_t win = 0;
        if (!j.desc.input_on_device)
            for (const ChunkIn &c : j.chunks)
                win = std::max<uint64_t>(win, (uint64_t)c.nf * (uint64_t)(info.bpf1 + info.bpf2));
        return 2ull * kNumProducts * (uint64_t)info.half * sizeof(double) + 2 * win;
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
        j.d_acc = static_cast<double *>(dev_get(acc_bytes, &j.acc_cap));
        j.d_chunk = static_cast<double *>(dev_get(acc_bytes, &j.chunk_cap));
        j.h_result = host_get(acc_bytes + sizeof(double), &j.res_cap);
        j.done_ev = ev_get();
        bool evs = j.done_ev != nullptr;
        if (!j.desc.input_on_device)
            for (int b = 0; b < 2; ++b) {
                j.copy_ev[b] = ev_get();
                j.comp_ev[b] = ev_get();
                evs = evs && j.copy_ev[b] && j.comp_ev[b];
            }
        if (!j.d_acc || !j.d_chunk || !j.h_result || !evs) {
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
        j.footprint = pb + j.acc_cap + j.chunk_cap;   // the input window is counted as it grows
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
                const int w = (int)(j.nchunk & 1);
                hipStream_t cs = copy_streams[j.stream];
                const size_t b[2] = {(size_t)c.nf * info.bpf1, (size_t)c.nf * info.bpf2};
                for (int i = 0; i < 2; ++i) {
                    if (b[i] > j.win_cap[w][i]) {   // grow: the chunk two back may still read it
                        if (j.win_used[w]) (void)hipEventSynchronize(j.comp_ev[w]);
                        dev_put(j.win[w][i], j.win_cap[w][i]);
                        active_bytes -= j.win_cap[w][i];
                        j.footprint -= j.win_cap[w][i];
                        j.win[w][i] = nullptr;
                        j.win_cap[w][i] = 0;
                        size_t cap = 0;
                        j.win[w][i] = dev_get(b[i], &cap);
                        if (!j.win[w][i]) {
                            fail_active(j, "input window allocation failed");
                            return true;
                        }
                        j.win_cap[w][i] = cap;
                        active_bytes += cap;
                        j.footprint += cap;
                        note_peak();
                    }
                }
                // the window's previous chunk has finished with it before the copy
                if ((j.win_used[w] && hipStreamWaitEvent(cs, j.comp_ev[w], 0) != hipSuccess) ||
                    hipMemcpyAsync(j.win[w][0], c.r1, b[0], hipMemcpyHostToDevice, cs) != hipSuccess ||
                    hipMemcpyAsync(j.win[w][1], c.r2, b[1], hipMemcpyHostToDevice, cs) != hipSuccess ||
                    hipEventRecord(j.copy_ev[w], cs) != hipSuccess ||
                    hipStreamWaitEvent(s, j.copy_ev[w], 0) != hipSuccess) {
                    fail_active(j, "input copy failed");
                    return true;
                }
                d1 = (const uint8_t *)j.win[w][0];
                d2 = (const uint8_t *)j.win[w][1];
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
            if (!j.desc.input_on_device) {
                const int w = (int)(j.nchunk & 1);
                if (hipEventRecord(j.comp_ev[w], s) != hipSuccess) {
                    fail_active(j, "event record failed");
                    return true;
                }
                j.win_used[w] = true;
            }
            j.nchunk++;
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
        (void)hipStreamSynchronize(copy_streams[j.stream]);
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
        hipStream_t st = nullptr, cs = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
            if (st) (void)hipStreamDestroy(st);
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            for (hipStream_t x : s->copy_streams) (void)hipStreamDestroy(x);
            delete s;
            return sfail(FXC_ERR_HIP, "stream creation failed");
        }
        s->streams.push_back(st);
        s->copy_streams.push_back(cs);
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
    s->release_pools();
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    for (hipStream_t st : s->copy_streams) (void)hipStreamDestroy(st);
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
