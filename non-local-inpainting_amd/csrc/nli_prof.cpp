// nli_prof.cpp -- optional per-kernel timing with HIP events recorded on the
// stream the kernels are launched on (the event pair brackets the launches).
#include "nli_internal.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace nli {
namespace prof {
namespace {

struct Rec {
    int kernel;
    hipStream_t stream;
    hipEvent_t a, b;
    double work, units;
    int launches;
};

struct State {
    std::mutex mu;
    std::atomic<unsigned> mask{0};       // bit k: kernel k is timed (read by every launching thread)
    std::atomic<unsigned> period{1};     // time one launch in `period` of each kernel
    std::atomic<unsigned long long> seen[TK_COUNT] = {};
    std::vector<Rec> open;           // begun, not ended (per kernel at most one)
    std::vector<Rec> done;
    std::vector<hipEvent_t> pool;
    int launches[TK_COUNT] = {};
    double ms[TK_COUNT] = {}, work[TK_COUNT] = {}, units[TK_COUNT] = {};
};

State &st()
{
    static State s;
    return s;
}

hipEvent_t take(State &s)
{
    if (!s.pool.empty()) {
        hipEvent_t e = s.pool.back();
        s.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    NLI_HIP(hipEventCreate(&e));
    return e;
}

void drain(State &s)
{
    for (Rec &r : s.done) {
        NLI_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        NLI_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        s.launches[r.kernel] += r.launches;
        s.ms[r.kernel] += ms;
        s.work[r.kernel] += r.work;
        s.units[r.kernel] += r.units;
        s.pool.push_back(r.a);
        s.pool.push_back(r.b);
    }
    s.done.clear();
}

}  // namespace

void enable(unsigned kernels, unsigned period)
{
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    drain(s);
    s.mask = kernels & ((1u << TK_COUNT) - 1);
    s.period = period > 0 ? period : 1;
    for (int k = 0; k < TK_COUNT; k++) {
        s.launches[k] = 0;
        s.ms[k] = s.work[k] = s.units[k] = 0;
        s.seen[k] = 0;
    }
}

// whether this thread's last begin() of a kernel was sampled (begin and end of a launch
// run on the thread that launches it); unsampled launches take no lock
thread_local bool t_sampled[TK_COUNT] = {};

void begin(int kernel, hipStream_t stream)
{
    State &s = st();
    if (!((s.mask.load(std::memory_order_relaxed) >> kernel) & 1u)) return;
    t_sampled[kernel] = s.seen[kernel].fetch_add(1, std::memory_order_relaxed) % s.period.load() == 0;
    if (!t_sampled[kernel]) return;
    std::lock_guard<std::mutex> g(s.mu);
    Rec r{kernel, stream, take(s), take(s), 0, 0, 0};
    NLI_HIP(hipEventRecord(r.a, stream));
    s.open.push_back(r);
}

void end(int kernel, hipStream_t stream, double work, double units, int launches)
{
    State &s = st();
    if (!((s.mask.load(std::memory_order_relaxed) >> kernel) & 1u) || !t_sampled[kernel]) return;
    t_sampled[kernel] = false;
    std::lock_guard<std::mutex> g(s.mu);
    for (size_t i = 0; i < s.open.size(); i++)
        if (s.open[i].kernel == kernel && s.open[i].stream == stream) {
            Rec r = s.open[i];
            s.open.erase(s.open.begin() + i);
            NLI_HIP(hipEventRecord(r.b, stream));
            r.work = work;
            r.units = units;
            r.launches = launches;
            s.done.push_back(r);
            if (s.done.size() > 4096) drain(s);
            return;
        }
}

void read(int kernel, int *launches, double *ms, double *work, double *units)
{
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    drain(s);
    if (kernel < 0 || kernel >= TK_COUNT) throw std::invalid_argument("unknown timed kernel");
    if (launches) *launches = s.launches[kernel];
    if (ms) *ms = s.ms[kernel];
    if (work) *work = s.work[kernel];
    if (units) *units = s.units[kernel];
}

}  // namespace prof
}  // namespace nli
