// serve.hip — the served request loop over docking pools (fa_serve_run).
//
// Host code only: requests are submitted to their receptor's pool as their
// arrival time passes, and every pool is stepped as soon as its own
// previous step has ended (a non-blocking query), so the pools' launches
// keep overlapping on the device and a slow pool does not hold the others
// back.  Finished jobs are collected when their pool's step has ended.
#include <algorithm>
#include <chrono>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "fa_ctx.h"

extern "C" {

int fa_serve_run(fa_pool* const* pools, int n_pools, const fa_request* req, int n_req, int realtime, fa_served* out,
                 fa_serve_stats* stats)
{
    if (!pools || n_pools < 1 || n_req < 0 || (n_req > 0 && (!req || !out)))
        return fa_fail(FA_ERR_INVALID, "bad serve arguments");
    for (int p = 0; p < n_pools; ++p)
        if (!pools[p]) return fa_fail(FA_ERR_INVALID, "null pool");
    for (int i = 0; i < n_req; ++i)
        if (req[i].pool < 0 || req[i].pool >= n_pools) return fa_fail(FA_ERR_INVALID, "request names no pool");
    using clock = std::chrono::steady_clock;
    std::vector<int> order(n_req);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].t < req[b].t; });
    std::map<std::pair<int, int64_t>, int> owner;  // (pool, job id) -> request
    std::vector<int> busy(n_pools, 0);
    std::vector<char> inflight(n_pools, 0);
    std::vector<fa_job_result> got(64);
    int nxt = 0, n_out = 0;
    int64_t steps = 0;
    uint64_t poses = 0;
    const auto t0 = clock::now();
    auto since = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
    auto any = [](const auto& v) { return std::any_of(v.begin(), v.end(), [](auto x) { return x != 0; }); };
    while (nxt < n_req || any(busy) || any(inflight)) {
        const double now = since();
        while (nxt < n_req && (!realtime || req[order[nxt]].t <= now)) {
            const int i = order[nxt];
            const int p = req[i].pool;
            int64_t id = -1;
            const int rc = fa_pool_submit(pools[p], &req[i].job, &id);
            if (rc != FA_SUCCESS) return rc;
            owner[{p, id}] = i;
            ++busy[p];
            ++nxt;
        }
        bool progressed = false;
        for (int p = 0; p < n_pools; ++p) {
            if (inflight[p]) {
                int b = 0;
                int rc = fa_pool_busy(pools[p], &b);
                if (rc != FA_SUCCESS) return rc;
                if (b) continue;
                int n = 0;
                for (;;) {  // a step may finish more jobs than the buffer holds
                    if ((rc = fa_pool_collect(pools[p], got.data(), (int)got.size(), &n)) == FA_SUCCESS) break;
                    if (got.size() >= (1u << 20)) return rc;
                    got.resize(got.size() * 2);
                }
                const double done = since();
                for (int q = 0; q < n; ++q) {
                    auto it = owner.find({p, got[q].id});
                    if (it == owner.end()) return fa_fail(FA_ERR_INVALID, "a pool returned a job it was not given");
                    const int i = it->second;
                    owner.erase(it);
                    fa_served& o = out[n_out++];
                    o.request = i;
                    o.latency_s = done - (realtime ? req[i].t : 0.0);
                    o.result = got[q];
                    poses += got[q].n_poses;
                    --busy[p];
                }
                inflight[p] = 0;
                progressed = true;
            }
            if (busy[p] > 0) {
                int running = 0, queued = 0;
                const int rc = fa_pool_step(pools[p], &running, &queued);
                if (rc != FA_SUCCESS) return rc;
                inflight[p] = 1;
                ++steps;
                progressed = true;
            }
        }
        if (progressed) continue;
        if (!any(inflight) && nxt < n_req) {
            const double wait = req[order[nxt]].t - since();
            if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        } else {
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
    if (stats) {
        stats->wall_s = since();
        stats->steps = steps;
        stats->poses = poses;
    }
    return FA_SUCCESS;
}

}  // extern "C"
