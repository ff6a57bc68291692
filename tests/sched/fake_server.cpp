// CPU build of the server's scheduling core (hera_gpu_amd/csrc/hv_sched.h)
// against a fake device, run under ThreadSanitizer and under
// AddressSanitizer + UndefinedBehaviorSanitizer by tests/test_sched_cpu.py.
//
// The fake executor gives each launch a thread that waits a few
// microseconds, writes the instance's memory and the slot's result buffer (a
// byte pattern from the job's seed), and reports completion or an injected
// device error.  Scenarios: a concurrent mix from several submitter threads
// with polls, waits and cancels from other threads; injected failures at
// admission, at launch and on the device; a job over the budget; a budget
// that forces instance eviction; shutdown with jobs still queued.
#include <atomic>
#include <cstdio>
#include <random>
#include <set>

#include "hv_sched.h"

static thread_local std::string g_err;
int hv::fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
extern "C" const char* hv_last_error(void) { return g_err.c_str(); }

using namespace hv::sched;

namespace {

struct FakeJob {
    uint32_t seed = 0;
    bool fail_launch = false, fail_device = false;
    int work_us = 0;
    size_t out_bytes = 0;
};

uint8_t pattern(uint32_t seed, size_t i) { return (uint8_t)((seed * 2654435761u + (uint32_t)i * 7u) >> 5); }

struct FakeInstance {
    std::vector<uint8_t> mem;
};

struct FakeSlot {
    std::vector<uint8_t> result;
    std::atomic<int> state{0};   // 0 idle, 1 running, 2 done, 3 device error
    std::thread th;
};

class FakeExecutor : public Executor {
public:
    explicit FakeExecutor(int n) : slots((size_t)n) {}
    ~FakeExecutor() override {
        for (auto& s : slots)
            if (s.th.joinable()) s.th.join();
    }
    int instance_create(const Job& j, void** out) override {
        auto* in = new FakeInstance;
        in->mem.assign(std::min<size_t>(j.footprint, 1 << 16), 0);
        *out = in;
        created++;
        return HV_OK;
    }
    void instance_destroy(void* h) override {
        delete static_cast<FakeInstance*>(h);
        destroyed++;
    }
    int launch(int slot, void* h, Job& j) override {
        const FakeJob& fj = *static_cast<const FakeJob*>(j.payload);
        if (fj.fail_launch) return hv::fail(HV_ERR_HIP, "injected launch failure");
        FakeSlot& s = slots[(size_t)slot];
        if (s.th.joinable()) s.th.join();
        s.result.assign(fj.out_bytes, 0);
        s.state.store(1);
        auto* in = static_cast<FakeInstance*>(h);
        const FakeJob copy = fj;
        s.th = std::thread([&s, in, copy] {
            std::this_thread::sleep_for(std::chrono::microseconds(copy.work_us));
            for (size_t i = 0; i < in->mem.size(); i += 61) in->mem[i] = (uint8_t)copy.seed;
            for (size_t i = 0; i < s.result.size(); ++i) s.result[i] = pattern(copy.seed, i);
            s.state.store(copy.fail_device ? 3 : 2);
        });
        return HV_OK;
    }
    int query(int slot) override {
        FakeSlot& s = slots[(size_t)slot];
        const int st = s.state.load();
        if (st == 1) return HV_ERR_BUSY;
        if (s.th.joinable()) s.th.join();
        if (st == 3) return hv::fail(HV_ERR_HIP, "injected device error");
        return HV_OK;
    }
    void drain(int slot) override {
        FakeSlot& s = slots[(size_t)slot];
        if (s.th.joinable()) s.th.join();
    }
    void* take_result(int slot) override {
        auto* r = new std::vector<uint8_t>();
        r->swap(slots[(size_t)slot].result);
        return r;
    }
    const void* result_data(void* r) override { return static_cast<std::vector<uint8_t>*>(r)->data(); }
    void release_result(void* r) override { delete static_cast<std::vector<uint8_t>*>(r); }
    void job_free(Job& j) override {
        delete static_cast<FakeJob*>(j.payload);
        j.payload = nullptr;
    }
    std::atomic<int64_t> created{0}, destroyed{0};

private:
    std::vector<FakeSlot> slots;
};

int failures = 0;
#define EXPECT(c, ...)                                   \
    do {                                                 \
        if (!(c)) {                                      \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);           \
            std::fprintf(stderr, "\n");                  \
            ++failures;                                  \
        }                                                \
    } while (0)

enum Kind { OK_JOB, FAIL_ADMIT, FAIL_LAUNCH, FAIL_DEVICE };

struct Spec {
    Kind kind = OK_JOB;
    int bucket = 0;
    uint32_t seed = 0;
    size_t out_bytes = 0;
    std::vector<uint8_t> out;
};

std::unique_ptr<Job> make_job(Spec& sp, size_t footprint, int work_us) {
    auto j = std::make_unique<Job>();
    j->key.mode = 1;
    j->key.nant = 37 + sp.bucket;
    j->key.kbucket = 1024;
    j->footprint = footprint;
    j->out_bytes = sp.out_bytes;
    j->out = sp.out.data();
    j->flops = 1.0;
    j->flags = sp.kind == FAIL_ADMIT ? HV_JOB_FAIL_TEST : 0;
    auto* fj = new FakeJob;
    fj->seed = sp.seed;
    fj->fail_launch = sp.kind == FAIL_LAUNCH;
    fj->fail_device = sp.kind == FAIL_DEVICE;
    fj->work_us = work_us;
    fj->out_bytes = sp.out_bytes;
    j->payload = fj;
    return j;
}

bool pattern_ok(const Spec& sp) {
    for (size_t i = 0; i < sp.out_bytes; ++i)
        if (sp.out[i] != pattern(sp.seed, i)) return false;
    return true;
}

// Concurrent mix: 3 submitter threads, a canceller, a poller and waits.
void scenario_mix(int njobs) {
    FakeExecutor ex(4);
    Server* s = create(&ex, 4, (size_t)6 << 20, 3);
    std::vector<Spec> specs((size_t)njobs);
    std::mt19937 rng(7);
    for (int i = 0; i < njobs; ++i) {
        Spec& sp = specs[(size_t)i];
        const int r = (int)(rng() % 20);
        sp.kind = r == 0 ? FAIL_ADMIT : r == 1 ? FAIL_LAUNCH : r == 2 ? FAIL_DEVICE : OK_JOB;
        sp.bucket = (int)(rng() % 6);
        sp.seed = rng();
        // a few outputs large enough to be pre-faulted by the touchers
        sp.out_bytes = (i % 37 == 0) ? (size_t)(9 << 20) : (size_t)(1 + rng() % 20000);
        sp.out.assign(sp.out_bytes, 0xAB);
    }
    std::vector<int64_t> ids((size_t)njobs, 0);
    std::vector<std::atomic<int>> submitted((size_t)njobs);
    std::atomic<int> next{0};
    auto submitter = [&] {
        std::mt19937 r2(std::hash<std::thread::id>()(std::this_thread::get_id()));
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= njobs) break;
            Spec& sp = specs[(size_t)i];
            const size_t fp = (size_t)(1 + sp.bucket) << 20;
            int64_t id = 0;
            const int rc = submit(s, make_job(sp, fp, (int)(r2() % 300)), &id);
            EXPECT(rc == HV_OK, "submit %d rc %d", i, rc);
            ids[(size_t)i] = id;
            submitted[(size_t)i].store(1);
            if (r2() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    };
    std::vector<std::thread> subs;
    for (int t = 0; t < 3; ++t) subs.emplace_back(submitter);
    // cancel every 11th job, whatever state it is in by then
    std::set<int64_t> cancelled;
    std::thread canceller([&] {
        for (int i = 0; i < njobs; i += 11) {
            while (!submitted[(size_t)i].load()) std::this_thread::yield();
            if (cancel(s, ids[(size_t)i]) == HV_OK) cancelled.insert(ids[(size_t)i]);
        }
    });
    for (auto& t : subs) t.join();
    canceller.join();
    // collect: poll until everything has been reported once
    std::unordered_map<int64_t, hv_job_info> infos;
    std::vector<hv_job_info> buf(64);
    while ((int)infos.size() < njobs) {
        int n = 0;
        EXPECT(poll(s, buf.data(), 64, &n) == HV_OK, "poll");
        for (int k = 0; k < n; ++k) {
            EXPECT(infos.count(buf[(size_t)k].id) == 0, "job %lld reported twice", (long long)buf[(size_t)k].id);
            infos[buf[(size_t)k].id] = buf[(size_t)k];
        }
        if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    for (int i = 0; i < njobs; ++i) {
        const Spec& sp = specs[(size_t)i];
        const hv_job_info& in = infos[ids[(size_t)i]];
        const bool was_cancelled = cancelled.count(ids[(size_t)i]) > 0;
        if (in.status == HV_ERR_CANCELED) {
            EXPECT(was_cancelled, "job %d cancelled without a request", i);
            continue;
        }
        if (sp.kind == FAIL_ADMIT) EXPECT(in.status == HV_ERR_STATE, "job %d status %d", i, in.status);
        else if (sp.kind == FAIL_LAUNCH || sp.kind == FAIL_DEVICE)
            EXPECT(in.status == HV_ERR_HIP, "job %d status %d", i, in.status);
        else {
            EXPECT(in.status == HV_OK, "job %d status %d (%s)", i, in.status, in.error);
            EXPECT(pattern_ok(sp), "job %d result differs", i);
        }
    }
    hv_server_stats_t st{};
    stats(s, &st);
    EXPECT(st.completed + st.failed == njobs, "completed %lld + failed %lld != %d", (long long)st.completed,
           (long long)st.failed, njobs);
    EXPECT(st.peak_running >= 2, "jobs never overlapped (peak %lld)", (long long)st.peak_running);
    EXPECT(st.workspace_bytes <= st.budget_bytes, "instances over budget");
    std::printf("mix: %d jobs, %lld completed, %lld failed, %lld cancelled, %lld instances, %lld evictions, "
                "peak %lld\n", njobs, (long long)st.completed, (long long)st.failed, (long long)st.cancelled,
                (long long)st.instances, (long long)st.evictions, (long long)st.peak_running);
    destroy(s);
    EXPECT(ex.created.load() == ex.destroyed.load(), "instances leaked: %lld created, %lld destroyed",
           (long long)ex.created.load(), (long long)ex.destroyed.load());
}

// A job over the budget is refused at submission; a budget of two instances
// with four buckets evicts idle instances and every job still completes.
void scenario_budget() {
    FakeExecutor ex(2);
    Server* s = create(&ex, 2, (size_t)2 << 20, 1);
    Spec big;
    big.out_bytes = 16;
    big.out.assign(16, 0);
    int64_t id = 0;
    auto jb = make_job(big, (size_t)3 << 20, 0);
    auto* payload = static_cast<FakeJob*>(jb->payload);
    const int rc = submit(s, std::move(jb), &id);
    EXPECT(rc == HV_ERR_NOMEM, "over-budget submit rc %d", rc);
    if (rc != HV_OK) delete payload;   // a refused job's payload stays with the caller
    std::vector<Spec> specs(40);
    std::vector<int64_t> ids;
    for (size_t i = 0; i < specs.size(); ++i) {
        specs[i].bucket = (int)(i % 4);
        specs[i].seed = (uint32_t)(1000 + i);
        specs[i].out_bytes = 4096;
        specs[i].out.assign(4096, 0);
        int64_t jid = 0;
        EXPECT(submit(s, make_job(specs[i], (size_t)1 << 20, 50), &jid) == HV_OK, "submit");
        ids.push_back(jid);
    }
    for (size_t i = 0; i < ids.size(); ++i) {
        hv_job_info in{};
        EXPECT(wait(s, ids[i], 30.0, &in) == HV_OK, "wait");
        EXPECT(in.status == HV_OK, "budget job %zu status %d", i, in.status);
        EXPECT(pattern_ok(specs[i]), "budget job %zu result differs", i);
    }
    hv_server_stats_t st{};
    stats(s, &st);
    EXPECT(st.evictions > 0, "no eviction with 4 buckets and room for 2 instances");
    EXPECT(st.instances <= 2, "%lld instances held", (long long)st.instances);
    EXPECT(wait(s, ids[0], 0.0, nullptr) == HV_ERR_ARG, "a collected job can be waited for again");
    EXPECT(cancel(s, ids[0]) == HV_ERR_ARG, "a collected job can be cancelled");
    std::printf("budget: %lld evictions, %lld instances\n", (long long)st.evictions, (long long)st.instances);
    destroy(s);
    EXPECT(ex.created.load() == ex.destroyed.load(), "instances leaked");
}

// Shutdown with queued, running and finished-but-uncollected jobs.
void scenario_shutdown() {
    FakeExecutor ex(2);
    Server* s = create(&ex, 2, (size_t)8 << 20, 2);
    std::vector<Spec> specs(60);
    for (size_t i = 0; i < specs.size(); ++i) {
        specs[i].bucket = (int)(i % 3);
        specs[i].seed = (uint32_t)i;
        specs[i].out_bytes = (i == 5) ? (size_t)(9 << 20) : 2048;
        specs[i].out.assign(specs[i].out_bytes, 0);
        int64_t jid = 0;
        EXPECT(submit(s, make_job(specs[i], (size_t)1 << 20, 500), &jid) == HV_OK, "submit");
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    destroy(s);
    EXPECT(ex.created.load() == ex.destroyed.load(), "instances leaked at shutdown");
    std::printf("shutdown: ok\n");
}

}  // namespace

int main(int argc, char** argv) {
    const int njobs = argc > 1 ? std::atoi(argv[1]) : 400;
    scenario_mix(njobs);
    scenario_budget();
    scenario_shutdown();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("all scheduling checks passed\n");
    return 0;
}
