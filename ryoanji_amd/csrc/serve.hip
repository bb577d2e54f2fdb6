// Serving: a continuous-batching request loop over resident particle systems.
//
// Requests (particle systems of any size, each running some leapfrog steps) arrive over time. A dispatcher
// admits an arrived request as soon as a worker slot is idle AND the request's HBM footprint fits in the
// remaining budget (first fit in arrival order). Every slot owns a HIP stream, a context and a resident
// system, so admitted requests run concurrently and their kernels overlap on the device; a request's
// particles, keys, tree and multipoles stay in HBM for all of its steps. Per request the loop records
// arrival, admission and completion times; with one slot it is the one-at-a-time baseline.
#include "common.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

namespace
{
using Clock = std::chrono::steady_clock;

size_t footprint(size_t n)
{
    // particles (2 x 9 arrays), keys x2, permutation x2, outputs, 32-byte gather records, and ~n/6 tree nodes x
    // ~230 B
    return n * (2 * 9 * 4 + 16 + 8 + 16 + 32) + (n / 6 + 64) * 230 + (64u << 10);
}
} // namespace

struct ryo_server
{
    int device = 0;
    int slots  = 1;
    size_t budget = 0;
    size_t maxN   = 0;
    struct Slot
    {
        hipStream_t stream = nullptr;
        ryo_ctx* ctx       = nullptr;
        ryo_system* sys    = nullptr;
        size_t cap         = 0;
    };
    std::vector<Slot> pool;
};

extern "C" {

int ryo_server_create(int slots, size_t hbm_budget, size_t max_n, ryo_server** out)
{
    if (!out || slots < 1 || slots > 64 || max_n == 0) return RYO_ERR_ARG;
    auto* s   = new ryo_server();
    s->slots  = slots;
    s->budget = hbm_budget ? hbm_budget : size_t(-1);
    s->maxN   = max_n;
    if (hipGetDevice(&s->device) != hipSuccess)
    {
        delete s;
        return RYO_ERR_HIP;
    }
    s->pool.resize(slots);
    for (auto& p : s->pool)
    {
        if (hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking) != hipSuccess ||
            ryo_ctx_create(&p.ctx, p.stream) != RYO_OK || ryo_system_create(p.ctx, max_n, &p.sys) != RYO_OK)
        {
            ryo_server_destroy(s);
            return RYO_ERR_HIP;
        }
        p.cap = max_n;
    }
    *out = s;
    return RYO_OK;
}

int ryo_server_destroy(ryo_server* s)
{
    if (!s) return RYO_OK;
    for (auto& p : s->pool)
    {
        if (p.sys) ryo_system_destroy(p.sys);
        if (p.ctx) ryo_ctx_destroy(p.ctx);
        if (p.stream) hipStreamDestroy(p.stream);
    }
    delete s;
    return RYO_OK;
}

size_t ryo_server_footprint(size_t n) { return footprint(n); }

/* Run a request stream to completion. Inputs and outputs are host memory. arrival_ms is relative to the start
 * of the run. Results: {arrival, admit, finish} ms and the final total potential per request; the final
 * state of request i (SFC order, n floats each of x, y, z, ax, ay, az, and the original index) is copied to
 * out_state[i] when out_state and out_state[i] are non-NULL (7*n floats). */
int ryo_server_run(ryo_server* s, int num, const ryo_request* req, ryo_result* res, float** out_state)
{
    if (!s || num < 0 || (num && (!req || !res))) return RYO_ERR_ARG;
    for (int i = 0; i < num; ++i)
        if (req[i].n == 0 || req[i].n > s->maxN || req[i].steps < 0) return RYO_ERR_ARG;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> waiting;
    std::vector<int> slotJob(s->slots, -1);
    size_t used = 0;
    int done = 0, next = 0;
    bool stop = false;
    int firstError = RYO_OK;
    const auto t0 = Clock::now();
    auto now = [&]() { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };

    auto worker = [&](int k) {
        hipSetDevice(s->device);
        auto& slot = s->pool[k];
        for (;;)
        {
            int j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || slotJob[k] >= 0; });
                if (slotJob[k] < 0) return;
                j = slotJob[k];
            }
            const ryo_request& r = req[j];
            ryo_params p;
            ryo_default_params(&p);
            p.theta = r.theta;
            p.G     = r.G;
            p.bucket_size = r.bucket_size ? r.bucket_size : 64;
            p.group_size  = r.group_size ? r.group_size : 64;
            double pot = 0;
            int rc = ryo_system_load(slot.sys, r.n, r.x, r.y, r.z, r.m, r.h, r.limits, r.periodic);
            if (!rc) rc = ryo_system_leapfrog(slot.sys, &p, r.dt, r.steps, &pot);
            double acc[4] = {0, 0, 0, 0};
            if (!rc && r.accuracy_targets > 0) rc = ryo_system_accuracy(slot.sys, r.accuracy_targets, uint64_t(j), acc);
            float lim[6] = {0, 0, 0, 0, 0, 0};
            if (!rc) rc = ryo_system_limits(slot.sys, lim);
            if (!rc && out_state && out_state[j])
            {
                float* a[8];
                float* o[4];
                int32_t* order;
                ryo_system_arrays(slot.sys, a);
                ryo_system_results(slot.sys, o, &order);
                const void* src[7] = {a[0], a[1], a[2], o[1], o[2], o[3], order};
                for (int q = 0; q < 7 && !rc; ++q)
                    if (hipMemcpyAsync(out_state[j] + q * r.n, src[q], r.n * 4, hipMemcpyDeviceToHost,
                                       slot.stream) != hipSuccess)
                        rc = RYO_ERR_HIP;
            }
            if (!rc && hipStreamSynchronize(slot.stream) != hipSuccess) rc = RYO_ERR_HIP;
            {
                std::lock_guard<std::mutex> lk(mu);
                res[j].finish_ms = now();
                res[j].total_pot = pot;
                res[j].err_median = acc[0], res[j].err_p99 = acc[1], res[j].err_max = acc[2];
                for (int q = 0; q < 6; ++q)
                    res[j].limits[q] = lim[q];
                res[j].status    = rc;
                res[j].slot      = k;
                if (rc && !firstError) firstError = rc;
                used -= footprint(r.n);
                slotJob[k] = -1;
                ++done;
            }
            cv.notify_all();
        }
    };

    std::vector<std::thread> threads;
    for (int k = 0; k < s->slots; ++k)
        threads.emplace_back(worker, k);

    {
        std::unique_lock<std::mutex> lk(mu);
        while (done < num)
        {
            double t = now();
            while (next < num && req[next].arrival_ms <= t)
            {
                res[next].arrival_ms = req[next].arrival_ms;
                waiting.push_back(next++);
            }
            // first fit in arrival order: an idle slot and room in the HBM budget
            bool admitted = true;
            while (admitted)
            {
                admitted = false;
                int idle = -1;
                for (int k = 0; k < s->slots; ++k)
                    if (slotJob[k] < 0)
                    {
                        idle = k;
                        break;
                    }
                if (idle < 0) break;
                for (auto it = waiting.begin(); it != waiting.end(); ++it)
                {
                    size_t f = footprint(req[*it].n);
                    if (used + f <= s->budget || used == 0)
                    {
                        used += f;
                        res[*it].admit_ms = now();
                        slotJob[idle]     = *it;
                        waiting.erase(it);
                        admitted = true;
                        break;
                    }
                }
            }
            cv.notify_all();
            if (done >= num) break;
            double wake = next < num ? req[next].arrival_ms - now() : 50.0;
            cv.wait_for(lk, std::chrono::duration<double, std::milli>(std::max(0.05, std::min(wake, 50.0))));
        }
        stop = true;
    }
    cv.notify_all();
    for (auto& th : threads)
        th.join();
    return firstError;
}

} // extern "C"
