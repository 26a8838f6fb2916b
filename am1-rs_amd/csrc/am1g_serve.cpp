// Continuous-batching SCF server (am1g_serve_*).
//
// Two ways of running requests. With `continuous` set (the default), one persistent service
// launch takes requests from a ring in pinned host memory as they are published: a request that
// arrives while others are on the device joins the running launch as soon as a workgroup is free,
// and computes its pair integrals, SCF and results there (the same arithmetic as a batched
// launch, so the same bits). The host thread of the service admits requests by device memory (a
// first-fit arena for each request's matrices, pair pool and DIIS history), publishes them,
// collects results as their done flags appear, and starts a new launch when the previous one has
// ended (its workgroups leave after a stretch without work). Requests beyond one workgroup's
// reach run on the grid-level path on a lane of their own.
//
// Without `continuous`, lanes:
// Each lane is one host thread that owns one am1g context (its own HIP stream and device
// pools). A free lane takes every queued request of the head request's precision that fits the
// device-memory budget, up to max_batch, and runs them as one persistent batched launch; other
// lanes keep admitting newly arrived requests meanwhile, so launches overlap on the device. Each
// molecule's result is handed out as soon as that molecule finishes (the launch streams results
// into pinned host memory), not when the whole launch ends. A molecule that am1g_scf would send
// to the grid-level path (more than the route_large_ao tunable's basis functions) runs alone
// through am1g_scf when it reaches the head of the queue, giving the result it would give alone,
// unless at least serve_large_batch such molecules (each within one workgroup's reach) are queued:
// then a lane runs all of them in one batched launch, which ends sooner than that many grid-path
// SCFs one after another. Requests are copied at submission, so callers may free their buffers
// at once.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "am1g_ctx.hpp"
#include "am1g_launch.hpp"

namespace {

using Clock = std::chrono::steady_clock;

enum class JobState { kQueued, kRunning, kDone };

struct Job {
    am1g_request rq;
    std::vector<uint8_t> z;
    std::vector<double> xyz;
    double field[3] = {0.0, 0.0, 0.0};  // the request's own copy of its external field
    size_t bytes = 0;
    JobState state = JobState::kQueued;
    bool cancelled = false;
    Clock::time_point submitted, done;
    am1g_result res;
    std::vector<double> charges, mo, mob, bcc;
    bool large = false;      // above the grid-path route: runs alone on the grid-level path ...
    bool batchable = false;  // ... unless enough such requests queue up (within one workgroup's reach)
};

// Device bytes one request holds while it runs: inside a batched launch, its dense matrices,
// pair pool, DIIS history and per-AO / per-atom arrays (the pool layout of am1g_scf_batch); on
// the grid-level path, about 13 dense matrices, the two packed DIIS histories of 8 entries each
// and the eigensolver's workspace (36 nao x nao doubles in all, beside the pair pool).
size_t job_bytes(const am1g_request& rq, bool large) {
    int64_t nao = 0, nh = 0, nH = 0;
    for (int a = 0; a < rq.n_atoms; ++a) {
        if (rq.z[a] == 1) {
            ++nao;
            ++nH;
        } else {
            nao += 4;
            ++nh;
        }
    }
    if (large)
        return (size_t)(36 * nao * nao + w_pool_size((int)nh, (int)nH) + (int64_t)rq.n_atoms * (rq.n_atoms + 8)) *
               sizeof(double);
    const bool uhf = rq.reference == AM1G_REF_UHF || rq.multiplicity > 1;
    const int64_t doubles = (int64_t)(uhf ? kMatSlotsUhf : kMatSlotsRhf) * nao * nao + w_pool_size((int)nh, (int)nH) +
                            hist_size((int)nao, uhf) + 2 * nao + (int64_t)rq.n_atoms * (rq.n_atoms + 8);
    return (size_t)doubles * sizeof(double) + sizeof(MolDesc) + sizeof(MolResult);
}

}  // namespace

// First-fit allocator over [0, size) in 256-byte granules, with coalescing on free.
class Arena {
  public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t& off) {
        const size_t need = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= need) {
                off = it->first;
                const size_t rest = it->second - need;
                free_.erase(it);
                if (rest) free_[off + need] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes) {
        const size_t len = (bytes + 255) & ~(size_t)255;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;  // offset -> length
};

// Device and pinned host state of the service launch.
struct Service {
    static constexpr int kCap = 1024;  // ring slots
    am1g_ctx* ctx = nullptr;           // stream, element tables, device pools
    DevBuf arena, z, aoff, acls, alist, xyz, mo, mob, charges, results, next, busy, last;
    HostBuf in, out, done, pub, stop, alive;
    hipStream_t extra = nullptr;       // a second stream: a new launch may join a partly emptied one
    Arena heap;
    size_t heap_bytes = 0;
    long long published = 0;           // tickets below this are in the ring
    long long slot_job[kCap];          // job ticket per slot, -1 when free
    size_t slot_off[kCap], slot_bytes[kCap];
    MolDesc slot_md[kCap];
    bool running[2] = {false, false};  // a service launch is on ctx->stream / extra
    bool paused = false;               // the grid lane holds the device: no new launch
    Clock::time_point started[2];
    BatchArgs args;
    ~Service() {
        if (extra) (void)hipStreamDestroy(extra);
    }
};

struct am1g_server {
    am1g_serve_options op;
    Service* svc = nullptr;
    std::thread svc_thread;
    std::vector<am1g_ctx*> ctxs;
    std::vector<std::thread> lanes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<long long> queue;
    std::unordered_map<long long, Job> jobs;
    long long next_ticket = 1;
    bool stop = false;
    size_t admitted = 0, peak = 0;
    am1g_serve_stats st{};
    std::string err;

    // Runs the queued request `t` alone through am1g_scf on `ctx`; called and returns with `lk` held.
    void run_large(std::unique_lock<std::mutex>& lk, am1g_ctx* ctx, long long t) {
        queue.erase(std::remove(queue.begin(), queue.end(), t), queue.end());
        Job& j = jobs.at(t);
        j.state = JobState::kRunning;
        const size_t bytes = j.bytes;
        admitted += bytes;
        peak = std::max(peak, admitted);
        st.launches++;
        st.queued = (long long)queue.size();
        st.running++;
        const am1g_request rq = j.rq;
        am1g_result res;
        std::memset(&res, 0, sizeof(res));
        res.charges = j.charges.data();
        res.mo_energies = j.mo.data();
        res.mo_energies_beta = j.mob.data();
        res.bcc_charges = j.bcc.empty() ? nullptr : j.bcc.data();
        lk.unlock();
        const int rc = am1g_scf(ctx, &rq, &res);
        const std::string msg = rc == AM1G_OK ? std::string() : std::string(am1g_last_error(ctx));
        const auto now = Clock::now();
        lk.lock();
        admitted -= bytes;
        auto it = jobs.find(t);  // cancelled jobs stay until polled, so the ticket is still here
        if (it != jobs.end()) {
            Job& d = it->second;
            d.res = res;
            if (rc != AM1G_OK && d.res.status == AM1G_OK) d.res.status = rc;
            d.state = JobState::kDone;
            d.done = now;
        }
        st.completed++;
        st.running--;
        if (rc != AM1G_OK) err = msg;
        cv_done.notify_all();
        cv_work.notify_all();
    }

    // The service launch's pools and pinned ring; called once at creation.
    int service_setup(am1g_ctx* ctx) {
        Service& v = *svc;
        v.ctx = ctx;
        constexpr int C = Service::kCap;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        v.heap_bytes = (size_t)(op.memory_budget_mb * 1048576.0);
        HIP_TRY(ctx, v.arena.reserve(v.heap_bytes));
        v.heap.reset(v.heap_bytes);
        HIP_TRY(ctx, v.z.reserve(sizeof(int) * C * kMaxAtoms));
        HIP_TRY(ctx, v.aoff.reserve(sizeof(int) * C * kMaxAtoms));
        HIP_TRY(ctx, v.acls.reserve(sizeof(int) * C * kMaxAtoms));
        HIP_TRY(ctx, v.alist.reserve(sizeof(int) * C * kMaxAtoms));
        HIP_TRY(ctx, v.xyz.reserve(sizeof(double) * C * 3 * kMaxAtoms));
        HIP_TRY(ctx, v.mo.reserve(sizeof(double) * C * kMaxBatchAo));
        HIP_TRY(ctx, v.mob.reserve(sizeof(double) * C * kMaxBatchAo));
        HIP_TRY(ctx, v.charges.reserve(sizeof(double) * C * kMaxAtoms));
        HIP_TRY(ctx, v.results.reserve(sizeof(MolResult) * C));
        HIP_TRY(ctx, v.next.reserve(sizeof(long long)));
        HIP_TRY(ctx, hipMemset(v.next.p, 0, sizeof(long long)));
        HIP_TRY(ctx, v.busy.reserve(sizeof(int)));
        HIP_TRY(ctx, hipMemset(v.busy.p, 0, sizeof(int)));
        HIP_TRY(ctx, v.last.reserve(sizeof(unsigned long long)));
        HIP_TRY(ctx, hipMemset(v.last.p, 0, sizeof(unsigned long long)));
        HIP_TRY(ctx, hipStreamCreateWithFlags(&v.extra, hipStreamNonBlocking));
        auto pinned = [&](HostBuf& b, size_t bytes) -> hipError_t {
            hipError_t e = hipHostMalloc(&b.p, bytes, hipHostMallocCoherent);
            if (e == hipSuccess) {
                b.bytes = bytes;
                std::memset(b.p, 0, bytes);
            }
            return e;
        };
        HIP_TRY(ctx, pinned(v.in, sizeof(SvcSlot) * C));
        HIP_TRY(ctx, pinned(v.out, sizeof(SvcOut) * C));
        HIP_TRY(ctx, pinned(v.done, sizeof(long long) * C));
        HIP_TRY(ctx, pinned(v.pub, sizeof(long long)));
        HIP_TRY(ctx, pinned(v.stop, sizeof(int)));
        HIP_TRY(ctx, pinned(v.alive, sizeof(int)));
        for (int k = 0; k < C; ++k) v.slot_job[k] = -1;
        BatchArgs& a = v.args;
        std::memset(&a, 0, sizeof(a));
        a.lds_n = ctx->tun.jacobi_lds_n;
        a.tri_max = std::min(eigen_tri_max(), ctx->tun.tri_max);
        a.z = v.z.as<int>();
        a.xyz = v.xyz.as<double>();
        a.elems = ctx->elems.as<DevElem>();
        a.mats = a.wpool = a.hist = v.arena.as<double>();
        a.mo_e = v.mo.as<double>();
        a.mo_e_beta = v.mob.as<double>();
        a.charges = v.charges.as<double>();
        a.results = v.results.as<MolResult>();
        a.aoff = v.aoff.as<int>();
        a.acls = v.acls.as<int>();
        a.alist = v.alist.as<int>();
        a.svc_in = v.in.as<SvcSlot>();
        a.svc_out = v.out.as<SvcOut>();
        a.svc_done = v.done.as<long long>();
        a.svc_published = v.pub.as<long long>();
        a.svc_stop = v.stop.as<int>();
        a.svc_next = v.next.as<long long>();
        a.svc_busy = v.busy.as<int>();
        a.svc_last = v.last.as<unsigned long long>();
        a.svc_alive = v.alive.as<int>();
        a.svc_cap = C;
        a.svc_idle_us = 20000;
        return AM1G_OK;
    }

    // Writes queued request t into ring slot k and publishes it; false when the device-memory
    // budget has no room for it now. Called with mu held.
    bool service_publish(long long t, std::string& msg, int& rc) {
        Service& v = *svc;
        constexpr int C = Service::kCap;
        const int k = (int)(v.published % C);
        Job& j = jobs.at(t);
        Prepared prep;
        rc = am1g_prepare(v.ctx, j.rq, prep, msg, kMaxBatchAo);
        if (rc != AM1G_OK) return true;  // refused: final at once
        MolDesc md = prep.md;
        const int64_t mats = (int64_t)(md.uhf ? kMatSlotsUhf : kMatSlotsRhf) * md.nao * md.nao;
        const int64_t wlen = std::max<int64_t>(w_pool_size(md.n_heavy, md.n_h), 1);
        const int64_t hlen = hist_size(md.nao, md.uhf);
        const size_t bytes = sizeof(double) * (size_t)(mats + wlen + hlen);
        size_t off = 0;
        if (!v.heap.alloc(bytes, off)) return false;
        const int64_t base = (int64_t)(off / sizeof(double));
        md.mat_off = base;
        md.w_off = base + mats;
        md.hist_off = base + mats + wlen;
        md.atom_off = (int64_t)k * kMaxAtoms;
        md.ao_off = (int64_t)k * kMaxBatchAo;
        SvcSlot& sl = v.in.as<SvcSlot>()[k];
        // heavy atoms first, then hydrogens, each in input order; AO offsets s,px,py,pz per heavy
        // atom and s per hydrogen in input order (basis.rs:37-60), as the batched launch builds them
        const DevElem* tab = v.ctx->host_tab + md.method * kMaxZ;
        int aoo = 0, nh = 0, nH = 0, nl = 0;
        for (int a = 0; a < j.rq.n_atoms; ++a) {
            const int zz = j.rq.z[a];
            sl.z[a] = zz;
            for (int q = 0; q < 3; ++q) sl.xyz[3 * a + q] = j.rq.xyz_bohr[3 * a + q];
            sl.aoff[a] = aoo;
            aoo += tab[zz].norb;
            sl.acls[a] = zz != 1 ? nh++ : nH++;
        }
        for (int a = 0; a < j.rq.n_atoms; ++a)
            if (j.rq.z[a] != 1) sl.alist[nl++] = a;
        for (int a = 0; a < j.rq.n_atoms; ++a)
            if (j.rq.z[a] == 1) sl.alist[nl++] = a;
        sl.md = md;
        v.slot_job[k] = t;
        v.slot_off[k] = off;
        v.slot_bytes[k] = bytes;
        v.slot_md[k] = md;
        j.state = JobState::kRunning;
        admitted += bytes;
        peak = std::max(peak, admitted);
        st.running++;
        ++v.published;
        __atomic_store_n(v.pub.as<long long>(), v.published, __ATOMIC_RELEASE);
        return true;
    }

    // Hands out every finished slot; called with mu held.
    void service_collect() {
        Service& v = *svc;
        constexpr int C = Service::kCap;
        const long long* done = v.done.as<long long>();
        const SvcOut* outs = v.out.as<SvcOut>();
        const long long lo = std::max(0LL, v.published - C);
        for (long long tk = lo; tk < v.published; ++tk) {
            const int k = (int)(tk % C);
            if (v.slot_job[k] < 0 || __atomic_load_n(done + k, __ATOMIC_ACQUIRE) != tk + 1) continue;
            const auto now = Clock::now();
            auto it = jobs.find(v.slot_job[k]);
            if (it != jobs.end()) {
                Job& j = it->second;
                const SvcOut& o = outs[k];
                const MolDesc& md = v.slot_md[k];
                am1g_result& r = j.res;
                std::memset(&r, 0, sizeof(r));
                am1g_fill_result(o.res, md, r);
                for (int q = 0; q < kPhases; ++q) r.phase_ms[q] = (double)o.res.phase[q] / 1.0e5;
                r.eigen_sweeps = o.res.eigen_sweeps;
                std::memcpy(j.charges.data(), o.charges, sizeof(double) * md.nat);
                std::memcpy(j.mo.data(), o.mo, sizeof(double) * md.nao);
                if (md.uhf) std::memcpy(j.mob.data(), o.mob, sizeof(double) * md.nao);
                r.charges = j.charges.data();
                r.bcc_charges = j.bcc.empty() ? nullptr : j.bcc.data();
                am1g_apply_bcc(j.rq, j.charges.data(), r);
                j.state = JobState::kDone;
                j.done = now;
            }
            v.heap.release(v.slot_off[k], v.slot_bytes[k]);
            admitted -= v.slot_bytes[k];
            v.slot_job[k] = -1;
            st.completed++;
            st.running--;
            cv_done.notify_all();
        }
    }

    void service_loop() {
        Service& v = *svc;
        constexpr int C = Service::kCap;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (stop) break;
            service_collect();
            // admit in arrival order while a ring slot and device memory are free
            bool blocked = false;
            for (auto it = queue.begin(); it != queue.end() && !blocked;) {
                Job& j = jobs.at(*it);
                if (!j.batchable) {  // the grid lane takes these
                    ++it;
                    continue;
                }
                if (v.slot_job[v.published % C] >= 0) break;  // the ring is full
                std::string msg;
                int rc = AM1G_OK;
                const long long t = *it;
                if (!service_publish(t, msg, rc)) {
                    blocked = true;
                    break;
                }
                it = queue.erase(it);
                if (rc != AM1G_OK) {
                    std::memset(&j.res, 0, sizeof(j.res));
                    j.res.status = rc;
                    j.state = JobState::kDone;
                    j.done = Clock::now();
                    err = msg;
                    st.completed++;
                    cv_done.notify_all();
                }
            }
            st.queued = (long long)queue.size();
            // A launch ends once its workgroups have all left (no request for a while and none
            // busy). Published work needs a launch: a new one when none runs, or a second one on
            // the other stream when most workgroups of the running one have already left (a
            // request came while they were leaving), so it gets the freed CUs.
            bool pending = false;
            for (int k = 0; k < C && !pending; ++k) pending = v.slot_job[k] >= 0;
            hipStream_t streams[2] = {v.ctx->stream, v.extra};
            for (int q = 0; q < 2; ++q) {
                if (!v.running[q]) continue;
                const hipError_t e = hipStreamQuery(streams[q]);
                if (e == hipSuccess) {
                    v.running[q] = false;
                    cv_work.notify_all();  // a paused grid lane waits for this
                } else if (e != hipErrorNotReady) {
                    err = std::string("HIP error in the service launch: ") + hipGetErrorString(e);
                    fail_all_running();
                    v.running[q] = false;
                }
            }
            const int alive = __atomic_load_n(v.alive.as<int>(), __ATOMIC_ACQUIRE);
            const int n_running = (int)v.running[0] + (int)v.running[1];
            const int q = v.running[0] ? 1 : 0;
            // (a launch's workgroups count themselves in as they start: look after a few ms)
            const bool thinned = n_running == 1 && 2 * alive < v.ctx->cus &&
                                 Clock::now() - v.started[1 - q] > std::chrono::milliseconds(5);
            if (pending && !v.paused && (n_running == 0 || thinned)) {
                const hipError_t e = launch_scf_batch(v.args, 1, v.ctx->cus, (size_t)kLdsBytes, streams[q]);
                if (e != hipSuccess) {
                    err = std::string("service launch failed: ") + hipGetErrorString(e);
                    fail_all_running();
                } else {
                    v.running[q] = true;
                    v.started[q] = Clock::now();
                    st.launches++;
                }
            }
            cv_work.wait_for(lk, std::chrono::microseconds(pending ? 50 : 1000));
        }
        // stop: the launch ends at its next check of the flag
        __atomic_store_n(v.stop.as<int>(), 1, __ATOMIC_RELEASE);
        lk.unlock();
        (void)hipStreamSynchronize(v.ctx->stream);
        (void)hipStreamSynchronize(v.extra);
    }

    void fail_all_running() {
        Service& v = *svc;
        for (int k = 0; k < Service::kCap; ++k) {
            if (v.slot_job[k] < 0) continue;
            auto it = jobs.find(v.slot_job[k]);
            if (it != jobs.end()) {
                it->second.res.status = AM1G_DEVICE_ERROR;
                it->second.state = JobState::kDone;
                it->second.done = Clock::now();
            }
            v.heap.release(v.slot_off[k], v.slot_bytes[k]);
            admitted -= v.slot_bytes[k];
            v.slot_job[k] = -1;
            st.completed++;
            st.running--;
        }
        cv_done.notify_all();
    }

    void lane_loop(int lane) {
        am1g_ctx* ctx = ctxs[lane];
        const size_t budget = (size_t)(op.memory_budget_mb * 1048576.0);
        // With two lanes or more, the first half of the lanes take only requests that batch
        // (at most route_large_ao basis functions), so a short request never waits for a lane
        // that is busy with long grid-path or large-molecule work; the other lanes take both.
        const bool small_only = !svc && op.lanes >= 2 && lane < op.lanes / 2;
        auto head_of = [&]() -> long long {
            for (long long q : queue) {
                const Job& j = jobs.at(q);
                if (svc ? !j.batchable : (!small_only || !j.large)) return q;
            }
            return -1;
        };
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] { return stop || head_of() >= 0; });
            if (stop) return;
            const long long head = head_of();
            // A grid-level request at the head runs alone, with no batching window, unless at
            // least serve_large_batch of them (of its precision, each within one workgroup's
            // reach) are queued: then one batched launch takes them all, which finishes sooner
            // than that many grid-path SCFs one after another.
            bool large_mode = false;
            if (jobs.at(head).large) {
                const Job& h = jobs.at(head);
                int n_large = 0;
                if (h.batchable)
                    for (long long q : queue) {
                        const Job& x = jobs.at(q);
                        n_large += x.large && x.batchable && x.rq.precision == h.rq.precision;
                    }
                large_mode = h.batchable && n_large >= ctx->tun.serve_large_batch;
            }
            if (svc) {
                // A grid-path request needs the whole device, which the service launch holds while it
                // has work: its workgroups leave after their current requests (the shared stop flag),
                // no new launch starts, and this lane runs every queued grid-path request before
                // handing the device back.
                Service& v = *svc;
                v.paused = true;
                __atomic_store_n(v.stop.as<int>(), 1, __ATOMIC_RELEASE);
                cv_work.wait(lk, [&] { return stop || (!v.running[0] && !v.running[1]); });
                if (!stop)
                    for (long long q = head_of(); q >= 0; q = head_of()) run_large(lk, ctx, q);
                if (stop) return;  // the device stop flag stays set: the server is closing
                __atomic_store_n(v.stop.as<int>(), 0, __ATOMIC_RELEASE);
                v.paused = false;
                cv_work.notify_all();
                continue;
            }
            if (jobs.at(head).large && !large_mode) {
                Job& j = jobs.at(head);
                const bool alone = admitted == 0;
                if (!alone && admitted + j.bytes > budget) {
                    cv_work.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
                run_large(lk, ctx, head);
                continue;
            }
            if (op.max_wait_ms > 0.0 && !large_mode) {  // batching window from the oldest queued request
                const auto until = jobs.at(head).submitted +
                                   std::chrono::duration_cast<Clock::duration>(
                                       std::chrono::duration<double, std::milli>(op.max_wait_ms));
                cv_work.wait_until(lk, until, [&] { return stop || (int)queue.size() >= op.max_batch; });
                if (stop) return;
                if (head_of() != head) continue;  // taken by another lane meanwhile
            }
            // admit: head request's precision, in arrival order, within the memory budget
            std::vector<long long> take;
            size_t bytes = 0;
            const int prec = jobs.at(head).rq.precision;
            for (auto it = queue.begin(); it != queue.end() && (int)take.size() < op.max_batch;) {
                Job& j = jobs.at(*it);
                if (j.rq.precision != prec || j.large != large_mode || (large_mode && !j.batchable)) {
                    ++it;
                    continue;
                }
                const bool alone = take.empty() && admitted == 0;  // an oversized request runs alone
                if (!alone && admitted + bytes + j.bytes > budget) break;
                take.push_back(*it);
                bytes += j.bytes;
                j.state = JobState::kRunning;
                it = queue.erase(it);
            }
            if (take.empty()) {  // memory is held by running launches: wait for one to finish
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            admitted += bytes;
            peak = std::max(peak, admitted);
            st.launches++;
            st.queued = (long long)queue.size();
            st.running += (long long)take.size();
            std::vector<am1g_request> rqs(take.size());
            std::vector<am1g_result> res(take.size());
            for (size_t k = 0; k < take.size(); ++k) {
                Job& j = jobs.at(take[k]);  // element references stay valid across rehashing
                rqs[k] = j.rq;
                std::memset(&res[k], 0, sizeof(am1g_result));
                res[k].charges = j.charges.data();
                res[k].mo_energies = j.mo.data();
                res[k].mo_energies_beta = j.mob.data();
                res[k].bcc_charges = j.bcc.empty() ? nullptr : j.bcc.data();
            }
            lk.unlock();
            // each molecule is handed out as soon as it finishes on the device
            auto on_done = [&](int k) {
                const auto now = Clock::now();
                std::lock_guard<std::mutex> g(mu);
                Job& j = jobs.at(take[k]);
                j.res = res[k];
                j.state = JobState::kDone;
                j.done = now;
                st.completed++;
                st.running--;
                cv_done.notify_all();
            };
            const int rc = am1g_scf_batch_stream(ctx, (int)take.size(), rqs.data(), res.data(), on_done);
            const std::string msg = rc == AM1G_OK ? std::string() : std::string(am1g_last_error(ctx));
            const auto now = Clock::now();
            lk.lock();
            admitted -= bytes;  // the lane's device pools are free again
            for (size_t k = 0; k < take.size(); ++k) {
                auto it = jobs.find(take[k]);
                if (it == jobs.end()) continue;  // handed out and already collected by poll
                Job& j = it->second;
                if (j.state == JobState::kDone) continue;
                j.res = res[k];
                j.res.status = rc != AM1G_OK ? rc : AM1G_DEVICE_ERROR;
                j.state = JobState::kDone;
                j.done = now;
                st.completed++;
                st.running--;
            }
            if (rc != AM1G_OK) err = msg;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

void am1g_serve_defaults(am1g_serve_options* op) {
    op->lanes = 4;
    op->max_batch = 512;
    op->memory_budget_mb = 32768.0;
    op->max_wait_ms = 0.0;
    op->continuous = 1;
}

int am1g_serve_create(int device, const am1g_serve_options* op, am1g_server** out) {
    if (!out) return AM1G_INVALID_INPUT;
    *out = nullptr;
    am1g_serve_options o;
    am1g_serve_defaults(&o);
    if (op) o = *op;
    if (o.lanes < 1 || o.lanes > 16 || o.max_batch < 1 || !(o.memory_budget_mb > 0.0) || !(o.max_wait_ms >= 0.0))
        return am1g_set_error(nullptr, AM1G_INVALID_INPUT,
                              "serving needs 1..16 lanes, max_batch >= 1, a positive memory budget and a "
                              "non-negative batching window");
    am1g_server* sv = new am1g_server();
    sv->op = o;
    const int n_ctx = o.continuous ? 2 : o.lanes;  // service + one grid lane, or the lanes
    for (int l = 0; l < n_ctx; ++l) {
        am1g_ctx* c = nullptr;
        const int rc = am1g_create(device, &c);
        if (rc != AM1G_OK) {
            for (am1g_ctx* x : sv->ctxs) am1g_destroy(x);
            delete sv;
            return rc;
        }
        sv->ctxs.push_back(c);
    }
    if (o.continuous) {
        sv->svc = new Service();
        const int rc = sv->service_setup(sv->ctxs[0]);
        if (rc != AM1G_OK) {
            for (am1g_ctx* x : sv->ctxs) am1g_destroy(x);
            delete sv->svc;
            delete sv;
            return rc;
        }
        sv->svc_thread = std::thread([sv] { sv->service_loop(); });
        sv->lanes.emplace_back([sv] { sv->lane_loop(1); });
    } else {
        for (int l = 0; l < o.lanes; ++l) sv->lanes.emplace_back([sv, l] { sv->lane_loop(l); });
    }
    *out = sv;
    return AM1G_OK;
}

void am1g_serve_destroy(am1g_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto& t : sv->lanes) t.join();
    if (sv->svc_thread.joinable()) sv->svc_thread.join();
    delete sv->svc;  // its buffers go before the context that owns the device
    for (am1g_ctx* c : sv->ctxs) am1g_destroy(c);
    delete sv;
}

int am1g_serve_submit(am1g_server* sv, const am1g_request* rq, long long* ticket) {
    if (!sv || !rq || !ticket) return AM1G_INVALID_INPUT;
    if (rq->n_atoms < 1 || !rq->z || !rq->xyz_bohr) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->err = "molecule has no atoms";
        return AM1G_INVALID_INPUT;
    }
    Job j;
    j.z.assign(rq->z, rq->z + rq->n_atoms);
    j.xyz.assign(rq->xyz_bohr, rq->xyz_bohr + 3 * (size_t)rq->n_atoms);
    j.rq = *rq;
    j.rq.z = j.z.data();
    j.rq.xyz_bohr = j.xyz.data();
    if (rq->electric_field)
        for (int k = 0; k < 3; ++k) j.field[k] = rq->electric_field[k];
    int nao = 0;
    for (uint8_t zz : j.z) nao += zz == 1 ? 1 : 4;
    j.large = nao > sv->ctxs[0]->tun.route_large_ao || rq->n_atoms > kMaxAtoms;
    j.batchable = nao <= kMaxBatchAo && rq->n_atoms <= kMaxAtoms;
    j.bytes = job_bytes(j.rq, j.large);
    j.charges.assign(rq->n_atoms, 0.0);
    if (rq->bcc) j.bcc.assign(rq->n_atoms, 0.0);
    j.mo.assign(nao, 0.0);
    j.mob.assign(nao, 0.0);
    j.submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return AM1G_INVALID_INPUT;
        const long long t = sv->next_ticket++;
        Job& slot = sv->jobs[t];
        slot = std::move(j);
        slot.rq.z = slot.z.data();  // the vectors moved with their heap buffers; re-point anyway
        slot.rq.xyz_bohr = slot.xyz.data();
        slot.rq.electric_field = rq->electric_field ? slot.field : nullptr;
        sv->queue.push_back(t);
        sv->st.submitted++;
        sv->st.queued = (long long)sv->queue.size();
        *ticket = t;
    }
    sv->cv_work.notify_all();  // lanes differ in what they take, so every idle lane looks
    return AM1G_OK;
}

int am1g_serve_poll(am1g_server* sv, long long ticket, double wait_ms, am1g_result* out) {
    if (!sv) return AM1G_INVALID_INPUT;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) {
        sv->err = "unknown ticket";
        return AM1G_INVALID_INPUT;
    }
    auto ready = [&] {
        auto f = sv->jobs.find(ticket);
        return f == sv->jobs.end() || f->second.state == JobState::kDone || f->second.cancelled;
    };
    if (!ready()) {
        if (wait_ms < 0) sv->cv_done.wait(lk, ready);
        else if (wait_ms > 0)
            sv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms), ready);
    }
    it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) {  // collected meanwhile by another caller
        sv->err = "unknown ticket";
        return AM1G_INVALID_INPUT;
    }
    Job& j = it->second;
    if (j.cancelled && j.state == JobState::kQueued) {
        sv->jobs.erase(ticket);
        return AM1G_CANCELLED;
    }
    if (j.state != JobState::kDone) return AM1G_PENDING;
    int status = j.cancelled ? AM1G_CANCELLED : j.res.status;
    if (out) {
        double* q = out->charges;
        double* mo = out->mo_energies;
        double* mob = out->mo_energies_beta;
        double* dens = out->density;
        double* coeff = out->mo_coeff;
        double* coeff_b = out->mo_coeff_beta;
        double* spin = out->spin_density;
        double* qbcc = out->bcc_charges;
        *out = j.res;
        out->charges = q;
        out->mo_energies = mo;
        out->mo_energies_beta = mob;
        out->density = dens;  // not kept by the server
        out->mo_coeff = coeff;
        out->mo_coeff_beta = coeff_b;
        out->spin_density = spin;
        out->bcc_charges = qbcc;
        out->latency_ms = std::chrono::duration<double, std::milli>(j.done - j.submitted).count();
        if (j.cancelled) out->status = AM1G_CANCELLED;
        if (q) std::memcpy(q, j.charges.data(), sizeof(double) * j.charges.size());
        if (qbcc && !j.bcc.empty()) std::memcpy(qbcc, j.bcc.data(), sizeof(double) * j.bcc.size());
        if (mo && j.res.nao > 0) std::memcpy(mo, j.mo.data(), sizeof(double) * j.res.nao);
        if (mob && j.res.nao > 0 && j.res.unrestricted) std::memcpy(mob, j.mob.data(), sizeof(double) * j.res.nao);
    }
    sv->jobs.erase(ticket);
    return status;
}

int am1g_serve_cancel(am1g_server* sv, long long ticket) {
    if (!sv) return AM1G_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) {
        sv->err = "unknown ticket";
        return AM1G_INVALID_INPUT;
    }
    Job& j = it->second;
    if (j.cancelled) return AM1G_OK;
    j.cancelled = true;
    sv->st.cancelled++;
    if (j.state == JobState::kQueued) {
        sv->queue.erase(std::remove(sv->queue.begin(), sv->queue.end(), ticket), sv->queue.end());
        sv->st.queued = (long long)sv->queue.size();
    }
    sv->cv_done.notify_all();
    return AM1G_OK;
}

int am1g_serve_get_stats(am1g_server* sv, am1g_serve_stats* out) {
    if (!sv || !out) return AM1G_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->st;
    out->queued = (long long)sv->queue.size();
    out->admitted_mb = (double)sv->admitted / 1048576.0;
    out->peak_admitted_mb = (double)sv->peak / 1048576.0;
    return AM1G_OK;
}

const char* am1g_serve_last_error(const am1g_server* sv) {
    static thread_local std::string copy;
    if (!sv) return "";
    am1g_server* m = const_cast<am1g_server*>(sv);
    std::lock_guard<std::mutex> lk(m->mu);
    copy = m->err;
    return copy.c_str();
}

}  // extern "C"
