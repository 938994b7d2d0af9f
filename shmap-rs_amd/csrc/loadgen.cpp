// Benchmark helper: replays an arrival schedule into a serving loop.  Each read
// is submitted when the clock reaches its arrival time (all reads due at the
// same moment go in one submission).  Used by the benchmark only.
#include <stdint.h>

#include <chrono>
#include <thread>
#include <vector>

#include "../../include/shmap_hip.h"

// Reads i = bases[off[i], off[i+1]) with tags i, due at t_begin + arrival_s[i]
// on the server's clock (arrival_s ascending), submitted without copies (the
// bases outlive the run).  Returns 0 or the first error.
extern "C" int shm_loadgen_replay(shm_server *s, const uint8_t *bases, const int64_t *off, const double *arrival_s,
                                  int32_t n, double t_begin) {
    std::vector<uint64_t> tags;
    std::vector<shm_read> reads;
    int32_t i = 0;
    while (i < n) {
        const double now = shm_server_now(s);
        const double due = t_begin + arrival_s[i];
        if (now < due) {
            const double wait = due - now;
            if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
            continue;
        }
        tags.clear();
        reads.clear();
        while (i < n && t_begin + arrival_s[i] <= now) {
            tags.push_back((uint64_t)i);
            reads.push_back(shm_read{bases + off[i], (int32_t)(off[i + 1] - off[i])});
            i++;
        }
        const int rc = shm_server_submit_borrowed(s, tags.data(), reads.data(), (int32_t)reads.size());
        if (rc) return rc;
    }
    return SHM_OK;
}
