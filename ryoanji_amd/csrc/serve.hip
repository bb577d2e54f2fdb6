// Serving: a continuous-batching request loop over resident particle systems.
//
// Requests (particle systems of any size, each running some leapfrog steps) arrive over time. A dispatcher
// admits an arrived request as soon as it has a place AND its HBM footprint fits in the remaining budget (first
// fit in arrival order, or least work first with a cap on the wait, ryo_server_set_order). Two kinds of place:
//   * a slot: a HIP stream, a context and a resident system of its own, for requests of any size; admitted
//     requests on different slots run concurrently and their kernels overlap on the device;
//   * the batch lane, for small requests (open or periodic): one stream, context and ryo_batch whose members are
//     packed one after the other and advanced in lockstep, one batched solve (every stage one launch sequence for
//     all members, per-system offset tables) per leapfrog step of all of them. Requests join at the next step
//     boundary and leave when their own steps are done, so the batch composition changes continuously.
// A request's particles, keys, tree and multipoles stay in HBM for all of its steps on either kind of place, and
// a lane member's trajectory is the same as on a slot (the same kick-drift arithmetic, box refit and solve).
// Per request the loop records arrival, admission and completion times; with one slot and no lane it is the
// one-at-a-time baseline.
#include "common.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <algorithm>
#include <array>
#include <tuple>

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

namespace ryo
{
int packedKickDrift(Context* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, const float* ax,
                    const float* ay, const float* az, size_t n, const uint32_t* P, int S, const float2* coef,
                    const PackedWrap* wrap);
int batchSetBoundaries(ryo_batch* b, const int* periodic, const int* shells, const int* ewald);
int packedBoxes(Context* ctx, const float* x, const float* y, const float* z, const uint32_t* P, int S, float* out);
void cubicBox(const float* mn, const float* mx, float* limits);
int sequence(Context* ctx, int32_t* out, size_t n, int32_t start);
int gather(Context* ctx, const int32_t* perm, size_t n, int num, const float* const* in, float* const* out);
int accuracyStats(Context* ctx, const float* x, const float* y, const float* z, const float* m, const float* h,
                  size_t n, const float* p, const float* ax, const float* ay, const float* az, double G, int shells,
                  const double* boxL, double ewaldL, int num_targets, uint64_t seed, double* out);
} // namespace ryo

using namespace ryo;

namespace
{
// the batch lane's packed state: particle arrays in the order of the last solve (each member in its own SFC order)
enum LaneArray : int
{
    kLX = 0, kLY, kLZ, kLM, kLH, kLVX, kLVY, kLVZ, kLO /* original index, int32 bits */, kLP, kLAX, kLAY, kLAZ,
    kLaneArrays
};

struct LaneMember
{
    int req         = -1;
    size_t n        = 0;
    bool solved     = false; // the initial solve is done (the member then takes kick-drift-kick steps)
    int stepsDone   = 0;
    float lim[6]    = {};
    double pot      = 0;
};

struct Lane
{
    size_t maxN = 0, capN = 0;
    int capS = 0;
    hipStream_t stream = nullptr;
    ryo_ctx* ctx       = nullptr;
    ryo_batch* batch   = nullptr;
    float* a[2][kLaneArrays] = {};
    int cur                  = 0;
    uint32_t* dP   = nullptr; // member offsets (S + 1)
    float2* dCoef  = nullptr; // per-member kick / drift coefficients
    float* dBox    = nullptr; // per-member extents [S][6]
    PackedWrap* dWrap = nullptr; // per-member periodic boxes

    int create(int device, size_t cn, int cs)
    {
        capN = cn, capS = cs;
        if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
            ryo_ctx_create(&ctx, stream) != RYO_OK || ryo_batch_create(ctx, capN, capS, &batch) != RYO_OK)
            return RYO_ERR_HIP;
        for (auto& set : a)
            for (auto& p : set)
                if (hipMalloc(&p, capN * sizeof(float)) != hipSuccess) return RYO_ERR_HIP;
        if (hipMalloc(&dP, (capS + 1) * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&dCoef, capS * sizeof(float2)) != hipSuccess ||
            hipMalloc(&dBox, capS * 6 * sizeof(float)) != hipSuccess ||
            hipMalloc(&dWrap, capS * sizeof(PackedWrap)) != hipSuccess)
            return RYO_ERR_HIP;
        return RYO_OK;
    }
    void destroy()
    {
        if (stream) hipStreamSynchronize(stream);
        for (auto& set : a)
            for (auto& p : set)
                if (p) hipFree(p), p = nullptr;
        for (void* p : {(void*)dP, (void*)dCoef, (void*)dBox, (void*)dWrap})
            if (p) hipFree(p);
        dP = nullptr, dCoef = nullptr, dBox = nullptr, dWrap = nullptr;
        if (batch) ryo_batch_destroy(batch), batch = nullptr;
        if (ctx) ryo_ctx_destroy(ctx), ctx = nullptr;
        if (stream) hipStreamDestroy(stream), stream = nullptr;
    }
};
struct LaneArrays
{
    const float* src[kLaneArrays];
    float* dst[kLaneArrays];
};

// dst member k's particles [dOff[k], dOff[k+1]) from src position sOff[k]: the remaining members moved together
__global__ __launch_bounds__(256) void laneCompactKernel(LaneArrays a, const uint32_t* __restrict__ dOff,
                                                        const uint32_t* __restrict__ sOff, int K)
{
    const uint32_t n = dOff[K];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const int k        = ryo::segmentOf(dOff, K, i);
        const uint32_t src = sOff[k] + (i - dOff[k]);
        for (int q = 0; q < kLaneArrays; ++q)
            a.dst[q][i] = a.src[q][src];
    }
}

int laneCompact(Lane& L, ryo::Context* c, const std::vector<uint32_t>& sOff, const std::vector<uint32_t>& dOff)
{
    const int K = int(sOff.size());
    // the offset tables go through the member-offset buffer (2K + 1 <= 2 capS words)
    std::vector<uint32_t> tab(dOff.begin(), dOff.end());
    tab.insert(tab.end(), sOff.begin(), sOff.end());
    uint32_t* dTab = reinterpret_cast<uint32_t*>(L.dBox); // 6 capS floats of room
    if (hipMemcpyAsync(dTab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, L.stream) != hipSuccess)
        return RYO_ERR_HIP;
    LaneArrays a;
    for (int q = 0; q < kLaneArrays; ++q)
        a.src[q] = L.a[L.cur][q], a.dst[q] = L.a[L.cur ^ 1][q];
    const uint32_t n = dOff.back();
    laneCompactKernel<<<std::max(1u, std::min(ryo::iceil(n, 256), 4096u)), 256, 0, L.stream>>>(a, dTab, dTab + K + 1,
                                                                                               K);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(L.stream) != hipSuccess) return RYO_ERR_HIP;
    L.cur ^= 1;
    return RYO_OK;
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
    // the batch lane (maxN 0: off); created on the first run that uses it
    size_t laneMaxN = size_t(1) << 16, laneCapN = size_t(1) << 22;
    int laneCapS    = 1024;
    Lane lane;
    bool laneReady = false;
    // admission order: 0 arrival; 1 least work (particles x solves) first, requests that have waited maxWaitMs
    // or longer (when it is positive) ahead of the others in arrival order
    int order        = 0;
    double maxWaitMs = 0.0;
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

int ryo_server_set_batching(ryo_server* s, size_t max_n, size_t capacity, int max_systems)
{
    if (!s || (max_n && (capacity < max_n || max_systems < 1))) return RYO_ERR_ARG;
    if (s->laneReady) s->lane.destroy(), s->laneReady = false;
    s->laneMaxN = max_n;
    s->laneCapN = capacity;
    s->laneCapS = max_systems;
    return RYO_OK;
}

int ryo_server_set_order(ryo_server* s, int order, double max_wait_ms)
{
    if (!s || order < 0 || order > 1 || !(max_wait_ms >= 0.0)) return RYO_ERR_ARG;
    s->order     = order;
    s->maxWaitMs = max_wait_ms;
    return RYO_OK;
}

int ryo_server_destroy(ryo_server* s)
{
    if (!s) return RYO_OK;
    if (s->laneReady) s->lane.destroy();
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

    // small requests with the default bucket and group size go to the batch lane (periodic ones with valid
    // settings and, for the Ewald correction, a cubic box; others take a slot, which reports their error)
    auto laneable = [&](const ryo_request& r) {
        const bool cubic = r.limits[1] - r.limits[0] == r.limits[3] - r.limits[2] &&
                           r.limits[3] - r.limits[2] == r.limits[5] - r.limits[4];
        const bool bounds = !r.periodic || (r.num_shells >= 0 && r.ewald >= 0 && r.ewald <= 2 && (!r.ewald || cubic));
        return s->laneMaxN > 0 && r.n <= s->laneMaxN && r.n <= s->laneCapN && bounds &&
               (r.bucket_size == 0 || r.bucket_size == 64) && (r.group_size == 0 || r.group_size == 64);
    };
    bool useLane = false;
    for (int i = 0; i < num; ++i)
        useLane |= laneable(req[i]);
    if (useLane && !s->laneReady)
    {
        if (s->lane.create(s->device, s->laneCapN, s->laneCapS) != RYO_OK)
        {
            s->lane.destroy();
            return RYO_ERR_HIP;
        }
        s->laneReady = true;
    }

    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> waiting, lanePending;
    std::vector<int> slotJob(s->slots, -1);
    size_t used = 0, laneN = 0;
    int laneCount = 0;
    int done = 0, next = 0;
    bool stop = false;
    int firstError = RYO_OK;
    const auto t0 = Clock::now();
    auto now = [&]() { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };

    // record a finished request (under the lock)
    auto record = [&](int j, int rc, int slot, double pot, const double* acc, const float* lim) {
        res[j].finish_ms = now();
        res[j].total_pot = pot;
        res[j].err_median = acc[0], res[j].err_p99 = acc[1], res[j].err_max = acc[2];
        for (int q = 0; q < 6; ++q)
            res[j].limits[q] = lim[q];
        res[j].status = rc;
        res[j].slot   = slot;
        if (rc && !firstError) firstError = rc;
        used -= footprint(req[j].n);
        ++done;
    };

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
            p.num_shells  = r.periodic ? r.num_shells : 0;
            p.ewald       = r.periodic ? r.ewald : 0;
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
                record(j, rc, k, pot, acc, lim);
                slotJob[k] = -1;
            }
            cv.notify_all();
        }
    };

    // the batch lane: members advance in lockstep, one batched solve per step of all of them
    auto laneWorker = [&]() {
        hipSetDevice(s->device);
        Lane& L      = s->lane;
        Context* c   = &L.ctx->c;
        hipStream_t st = L.stream;
        std::vector<LaneMember> mem;
        size_t N = 0;
        for (;;)
        {
            std::vector<int> joining;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !lanePending.empty() || !mem.empty(); });
                if (mem.empty() && lanePending.empty()) return;
                while (!lanePending.empty())
                    joining.push_back(lanePending.front()), lanePending.pop_front();
            }
            int rc = RYO_OK;
            float** A = L.a[L.cur];
            // members join with zero velocity and acceleration and their requests' boxes
            for (int j : joining)
            {
                const ryo_request& r = req[j];
                LaneMember m;
                m.req = j, m.n = r.n;
                std::memcpy(m.lim, r.limits, sizeof(m.lim));
                const float* src[5] = {r.x, r.y, r.z, r.m, r.h};
                for (int k = 0; k < 5 && !rc; ++k)
                    if (hipMemcpyAsync(A[kLX + k] + N, src[k], r.n * 4, hipMemcpyDefault, st) != hipSuccess)
                        rc = RYO_ERR_HIP;
                for (int k : {int(kLVX), int(kLVY), int(kLVZ), int(kLP), int(kLAX), int(kLAY), int(kLAZ)})
                    if (!rc && hipMemsetAsync(A[k] + N, 0, r.n * 4, st) != hipSuccess) rc = RYO_ERR_HIP;
                if (!rc) rc = sequence(c, reinterpret_cast<int32_t*>(A[kLO]) + N, r.n, 0);
                N += r.n;
                mem.push_back(m);
            }
            const int S = int(mem.size());
            std::vector<uint32_t> hP(S + 1, 0);
            std::vector<float2> coef(S);
            std::vector<float> hBox(6 * size_t(S)), limits(6 * size_t(S)), theta(S), G(S);
            std::vector<uint32_t> counts(S);
            std::vector<double> pots(S, 0.0);
            std::vector<int> per(S), shells(S), ewm(S);
            std::vector<PackedWrap> wrap(S);
            bool anyPeriodic = false;
            for (int k = 0; k < S; ++k)
            {
                const ryo_request& r = req[mem[k].req];
                hP[k + 1] = hP[k] + uint32_t(mem[k].n);
                counts[k] = uint32_t(mem[k].n);
                coef[k]   = mem[k].solved ? make_float2(0.5f * r.dt, r.dt) : make_float2(0.0f, 0.0f);
                theta[k] = r.theta, G[k] = r.G;
                // a periodic member keeps its box and wraps into it, as a resident periodic system does
                per[k] = r.periodic != 0, shells[k] = per[k] ? r.num_shells : 0, ewm[k] = per[k] ? r.ewald : 0;
                anyPeriodic |= per[k] != 0;
                const float* l = r.limits;
                wrap[k] = PackedWrap{{l[0], l[2], l[4]}, {l[1] - l[0], l[3] - l[2], l[5] - l[4]}, per[k], 0};
            }
            // kick and drift of the members past their initial solve, then the open members' boxes refit around
            // them
            if (!rc && (hipMemcpyAsync(L.dP, hP.data(), hP.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
                        hipMemcpyAsync(L.dCoef, coef.data(), coef.size() * sizeof(float2), hipMemcpyHostToDevice,
                                       st) != hipSuccess ||
                        (anyPeriodic && hipMemcpyAsync(L.dWrap, wrap.data(), wrap.size() * sizeof(PackedWrap),
                                                       hipMemcpyHostToDevice, st) != hipSuccess)))
                rc = RYO_ERR_HIP;
            const PackedWrap* dWrap = anyPeriodic ? L.dWrap : nullptr;
            if (!rc) rc = packedKickDrift(c, A[kLX], A[kLY], A[kLZ], A[kLVX], A[kLVY], A[kLVZ], A[kLAX], A[kLAY],
                                          A[kLAZ], N, L.dP, S, L.dCoef, dWrap);
            if (!rc) rc = packedBoxes(c, A[kLX], A[kLY], A[kLZ], L.dP, S, L.dBox);
            if (!rc && (hipMemcpyAsync(hBox.data(), L.dBox, hBox.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess))
                rc = RYO_ERR_HIP;
            for (int k = 0; k < S; ++k)
            {
                if (mem[k].solved && !per[k]) cubicBox(&hBox[6 * k], &hBox[6 * k + 3], mem[k].lim);
                std::memcpy(&limits[6 * k], mem[k].lim, sizeof(mem[k].lim));
            }
            // one batched solve of every member
            ryo_params prm;
            ryo_default_params(&prm);
            if (!rc) rc = ryo_batch_load(L.batch, S, counts.data(), A[kLX], A[kLY], A[kLZ], A[kLM], A[kLH],
                                         limits.data());
            if (!rc && anyPeriodic) rc = batchSetBoundaries(L.batch, per.data(), shells.data(), ewm.data());
            if (!rc) rc = ryo_batch_solve(L.batch, &prm, theta.data(), G.data(), pots.data());
            float* out[9];
            int32_t* order = nullptr;
            if (!rc) rc = ryo_batch_results(L.batch, out, &order);
            float** B = L.a[L.cur ^ 1];
            if (!rc)
            {
                // the solve's SFC order: positions, masses, softenings and results from the batch, velocities
                // and original indices permuted the same way
                const int dst[9] = {kLX, kLY, kLZ, kLM, kLH, kLP, kLAX, kLAY, kLAZ};
                for (int k = 0; k < 9 && !rc; ++k)
                    if (hipMemcpyAsync(B[dst[k]], out[k], N * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
                        rc = RYO_ERR_HIP;
                const float* gin[4] = {A[kLVX], A[kLVY], A[kLVZ], A[kLO]};
                float* gout[4]      = {B[kLVX], B[kLVY], B[kLVZ], B[kLO]};
                if (!rc) rc = gather(c, order, N, 4, gin, gout);
                L.cur ^= 1;
                A = B;
            }
            // the closing half kick of the members that took a step
            for (int k = 0; k < S; ++k)
                coef[k] = mem[k].solved ? make_float2(0.5f * req[mem[k].req].dt, 0.0f) : make_float2(0.0f, 0.0f);
            if (!rc && hipMemcpyAsync(L.dCoef, coef.data(), coef.size() * sizeof(float2), hipMemcpyHostToDevice,
                                      st) != hipSuccess)
                rc = RYO_ERR_HIP;
            if (!rc) rc = packedKickDrift(c, A[kLX], A[kLY], A[kLZ], A[kLVX], A[kLVY], A[kLVZ], A[kLAX], A[kLAY],
                                          A[kLAZ], N, L.dP, S, L.dCoef, dWrap);
            // members whose steps are done leave with their final state
            std::vector<char> finished(S, 0);
            std::vector<std::array<double, 4>> accs(S, std::array<double, 4>{0, 0, 0, 0});
            for (int k = 0; k < S; ++k)
            {
                LaneMember& m = mem[k];
                if (m.solved) ++m.stepsDone;
                m.solved = true;
                m.pot    = pots[k];
                const ryo_request& r = req[m.req];
                finished[k] = rc != RYO_OK || m.stepsDone >= r.steps;
                if (!finished[k] || rc) continue;
                const size_t o = hP[k];
                if (r.accuracy_targets > 0)
                {
                    // as ryo_system_accuracy: the direct sum over the traversal's images, or the Ewald pair sum
                    // after an Ewald-corrected solve
                    const double boxL[3] = {double(m.lim[1] - m.lim[0]), double(m.lim[3] - m.lim[2]),
                                            double(m.lim[5] - m.lim[4])};
                    const double ewaldL  = per[k] && ewm[k] ? boxL[0] : 0.0;
                    rc = accuracyStats(c, A[kLX] + o, A[kLY] + o, A[kLZ] + o, A[kLM] + o, A[kLH] + o, m.n, A[kLP] + o,
                                       A[kLAX] + o, A[kLAY] + o, A[kLAZ] + o, double(r.G), shells[k], boxL, ewaldL,
                                       r.accuracy_targets, uint64_t(m.req), accs[k].data());
                }
                if (!rc && out_state && out_state[m.req])
                {
                    const int src[7] = {kLX, kLY, kLZ, kLAX, kLAY, kLAZ, kLO};
                    for (int q = 0; q < 7 && !rc; ++q)
                        if (hipMemcpyAsync(out_state[m.req] + q * m.n, A[src[q]] + o, m.n * 4, hipMemcpyDeviceToHost,
                                           st) != hipSuccess)
                            rc = RYO_ERR_HIP;
                }
            }
            if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = RYO_ERR_HIP;
            // the remaining members close ranks (one copy kernel over the 13 arrays)
            std::vector<LaneMember> keep;
            std::vector<uint32_t> sOff, dOff{0};
            for (int k = 0; k < S; ++k)
                if (!finished[k] && !rc)
                    keep.push_back(mem[k]), sOff.push_back(hP[k]), dOff.push_back(dOff.back() + uint32_t(mem[k].n));
            if (!rc && !keep.empty() && keep.size() < mem.size())
                rc = laneCompact(L, c, sOff, dOff);
            {
                std::lock_guard<std::mutex> lk(mu);
                for (int k = 0; k < S; ++k)
                    if (finished[k] || rc)
                    {
                        record(mem[k].req, rc, s->slots, mem[k].pot, accs[k].data(), mem[k].lim);
                        laneN -= mem[k].n;
                        --laneCount;
                    }
            }
            if (rc) keep.clear();
            mem.swap(keep);
            N = dOff.back();
            if (rc) N = 0;
            cv.notify_all();
        }
    };

    std::vector<std::thread> threads;
    for (int k = 0; k < s->slots; ++k)
        threads.emplace_back(worker, k);
    if (useLane) threads.emplace_back(laneWorker);

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
            if (s->order == 1 && waiting.size() > 1)
            {
                // least work first (the initial solve and one per step, times the particles); with a positive
                // maxWaitMs, requests that have waited that long go ahead of the others, in arrival order
                const double tNow = now();
                auto key = [&](int j) {
                    const bool aged = s->maxWaitMs > 0.0 && tNow - req[j].arrival_ms >= s->maxWaitMs;
                    return std::make_tuple(aged ? 0 : 1, aged ? 0.0 : double(req[j].n) * double(req[j].steps + 1), j);
                };
                std::sort(waiting.begin(), waiting.end(), [&](int a, int b) { return key(a) < key(b); });
            }
            // first fit in the admission order: the lane for a small request while it has room, else an idle slot;
            // both within the HBM budget
            for (auto it = waiting.begin(); it != waiting.end();)
            {
                const int j    = *it;
                const size_t f = footprint(req[j].n);
                const bool fits = used + f <= s->budget || used == 0;
                if (fits && useLane && laneable(req[j]) && laneN + req[j].n <= s->laneCapN &&
                    laneCount < s->laneCapS)
                {
                    used += f;
                    laneN += req[j].n;
                    ++laneCount;
                    res[j].admit_ms = now();
                    lanePending.push_back(j);
                    it = waiting.erase(it);
                    continue;
                }
                int idle = -1;
                for (int k = 0; k < s->slots && idle < 0; ++k)
                    if (slotJob[k] < 0) idle = k;
                if (fits && idle >= 0)
                {
                    used += f;
                    res[j].admit_ms = now();
                    slotJob[idle]   = j;
                    it = waiting.erase(it);
                    continue;
                }
                ++it;
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
