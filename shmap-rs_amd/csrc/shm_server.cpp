// Serving loop over one resident index: a request queue, and worker threads
// that each own a mapper (and so a HIP stream) and map whatever is queued, up
// to an admission budget, as one batch.  Built only on the public C ABI.
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/shmap_hip.h"

namespace {

struct Request {
    uint64_t tag;
    double submit_s;
    const uint8_t *seq;         // owned.data(), or the caller's bytes when borrowed
    int32_t len;
    std::vector<uint8_t> owned;
};

}  // namespace

struct shm_server {
    shm_server_config cfg{};
    std::vector<shm_mapper *> mappers;
    std::vector<std::thread> threads;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::deque<shm_served> done;
    int64_t running = 0;
    int64_t bstats[4] = {};  // batches, redone, allocating, captures
    bool stop = false;
    std::string error;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void worker(int wi);
};

void shm_server::worker(int wi) {
    shm_mapper *mp = mappers[wi];
    std::vector<Request> batch;
    std::vector<shm_read> reads;
    std::vector<shm_read_result> res;
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop) return;
            int64_t bases = 0;
            while (!queue.empty() && (int64_t)batch.size() < cfg.max_batch_reads) {
                const int64_t len = queue.front().len;
                if (!batch.empty() && bases + len > cfg.max_batch_bases) break;
                bases += len;
                batch.push_back(std::move(queue.front()));
                queue.pop_front();
            }
            running += (int64_t)batch.size();
        }
        const double start = now();
        const int32_t n = (int32_t)batch.size();
        reads.resize(n);
        res.assign(n, shm_read_result{});
        for (int32_t i = 0; i < n; i++) reads[i] = shm_read{batch[i].seq, batch[i].len};
        // a read that fails does so alone (its record's status); the others
        // of the batch are mapped and delivered as usual
        const int rc = shm_map_batch(mp, reads.data(), n, res.data());
        const std::string msg = rc ? shm_last_error() : "";
        const double end = now();
        int64_t wc[14] = {};
        int32_t nw = 0;
        (void)shm_mapper_last_counts(mp, wc, 14, &nw);
        bool marked = false;
        for (int32_t i = 0; i < n && !marked; i++) marked = res[i].status != 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (rc && error.empty()) error = msg;
            bstats[0]++;
            bstats[1] += wc[10] != 0;
            bstats[2] += wc[11] != 0;
            bstats[3] += wc[13];
            for (int32_t i = 0; i < n; i++) {
                shm_served d{};
                d.tag = batch[i].tag;
                d.submit_s = batch[i].submit_s;
                d.start_s = start;
                d.done_s = end;
                d.batch_reads = n;
                d.worker = wi;
                d.r = res[i];
                if (rc && !marked) d.r.status = rc;  // a failure before any read was looked at
                done.push_back(d);
            }
            running -= n;
        }
        cv_done.notify_all();
    }
}

extern "C" int shm_server_create(const shm_index *idx, const shm_map_params *params, const shm_server_config *cfg,
                                 shm_server **out) {
    if (!idx || !params || !out) return SHM_ERR_ARG;
    *out = nullptr;
    shm_server *s = new shm_server;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.workers <= 0) s->cfg.workers = 3;  // three streams overlap best on one device
    if (s->cfg.max_batch_reads <= 0) s->cfg.max_batch_reads = 65536;
    if (s->cfg.max_batch_bases <= 0) s->cfg.max_batch_bases = (int64_t)1 << 31;
    for (int i = 0; i < s->cfg.workers; i++) {
        shm_mapper *mp = nullptr;
        const int rc = shm_mapper_create(idx, params, &mp);
        if (rc) {
            for (shm_mapper *m : s->mappers) shm_mapper_free(m);
            delete s;
            return rc;
        }
        s->mappers.push_back(mp);
        if (s->cfg.max_read_len > 0) (void)shm_mapper_set_read_limit(mp, s->cfg.max_read_len);
        if (s->cfg.flags & SHM_SERVER_NO_SPECULATION) (void)shm_mapper_set_speculation(mp, 0);
        if (s->cfg.flags & SHM_SERVER_NO_GRAPHS) (void)shm_mapper_set_graphs(mp, 0);
    }
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < s->cfg.workers; i++) s->threads.emplace_back(&shm_server::worker, s, i);
    *out = s;
    return SHM_OK;
}

static int submit(shm_server *s, const uint64_t *tags, const shm_read *reads, int32_t n, bool borrow) {
    if (!s || n < 0 || (n > 0 && (!tags || !reads))) return SHM_ERR_ARG;
    for (int32_t i = 0; i < n; i++)
        if (reads[i].len < 0 || (reads[i].len > 0 && !reads[i].seq)) return SHM_ERR_ARG;
    std::vector<Request> rq(n);
    for (int32_t i = 0; i < n; i++) {
        rq[i].tag = tags[i];
        rq[i].len = reads[i].len;
        if (borrow) {
            rq[i].seq = reads[i].seq;
        } else {
            rq[i].owned.assign(reads[i].seq, reads[i].seq + reads[i].len);
            rq[i].seq = rq[i].owned.data();
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = s->now();
        for (Request &r : rq) {
            r.submit_s = t;
            s->queue.push_back(std::move(r));
        }
    }
    s->cv_work.notify_all();
    return SHM_OK;
}

extern "C" int shm_server_submit_many(shm_server *s, const uint64_t *tags, const shm_read *reads, int32_t n) {
    return submit(s, tags, reads, n, false);
}

extern "C" int shm_server_submit_borrowed(shm_server *s, const uint64_t *tags, const shm_read *reads, int32_t n) {
    return submit(s, tags, reads, n, true);
}

extern "C" int shm_server_submit(shm_server *s, uint64_t tag, const uint8_t *seq, int32_t len) {
    const shm_read r{seq, len};
    return shm_server_submit_many(s, &tag, &r, 1);
}

extern "C" int shm_server_poll(shm_server *s, shm_served *out, int32_t cap, int32_t min_results, int32_t timeout_ms,
                               int32_t *n) {
    if (!s || !n || cap < 0 || (cap > 0 && !out)) return SHM_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
        return (int64_t)s->done.size() >= min_results || (s->queue.empty() && s->running == 0);
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    int32_t k = 0;
    while (k < cap && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return SHM_OK;
}

extern "C" int shm_server_counts(shm_server *s, int64_t *queued, int64_t *running, int64_t *finished) {
    if (!s) return SHM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int64_t)s->queue.size();
    if (running) *running = s->running;
    if (finished) *finished = (int64_t)s->done.size();
    return SHM_OK;
}

extern "C" int shm_server_batch_stats(shm_server *s, int64_t *stats, int32_t n) {
    if (!s || n < 0 || (n > 0 && !stats)) return SHM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    for (int32_t i = 0; i < n && i < 4; i++) stats[i] = s->bstats[i];
    return SHM_OK;
}

extern "C" double shm_server_now(const shm_server *s) { return s ? s->now() : 0.0; }

extern "C" const char *shm_server_error(shm_server *s) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    thread_local std::string copy;
    copy = s->error;
    return copy.c_str();
}

extern "C" void shm_server_free(shm_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->threads) t.join();
    for (shm_mapper *m : s->mappers) shm_mapper_free(m);
    delete s;
}
