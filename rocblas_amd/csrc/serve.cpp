// serve.cpp -- continuous-batching serving loop for triangular solves.
//
// A server owns one device arena (the HBM admission budget) and a few
// streams, each with its own rocblas_handle.  rocblas_amd_server_run replays a
// trace of solve requests (arrival time, factor id, size, precision, seeds):
//
//   * arrival: a request joins the pending queue once its arrival time has
//     passed on the host clock;
//   * admission (FIFO): a request is admitted when its right-hand side, and
//     its factor unless the factor is already resident, fit in the arena;
//     resident factors that no in-flight request uses are evicted
//     least-recently-used first to make room; at most `max_inflight` requests
//     are on the device at once;
//   * launch: admitted requests are generated in place (factor on a miss,
//     right-hand side always) and solved: those with k <= `small_k` go
//     together into one rocblas_amd_{s,d}trsm_vbatched launch per precision on
//     stream 0, larger ones each into rocblas_{s,d}trsm rotating over the
//     other streams, so small requests never queue behind large solves and
//     work admitted while other work runs overlaps on the device;
//   * completion: an event per launch is polled; a request's latency is the
//     host time from its arrival to the poll that saw its launch complete.
//     Its right-hand side buffer is freed (or kept for
//     rocblas_amd_server_result) and its factor stays resident.
//
// Requests are side left, lower, no transpose, non-unit, alpha = 1, with the
// factor drawn from rocblas_amd_{s,d}synth_tri (uniform) under the factor's
// seed and the right-hand side from rocblas_amd_{s,d}synth_rhs under the
// request's seed, so a request's result does not depend on what else was in
// flight: every kernel it goes through treats instances independently.
#include "internal.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <set>
#include <thread>
#include <unordered_map>
#include <vector>

namespace
{
using clock_type = std::chrono::steady_clock;

// Best-fit sub-allocator over one device allocation (offsets, 256-byte units).
// Free blocks are indexed by offset (coalescing on release) and by size, so an
// allocation and the "does a hole of this size exist" test are O(log n) however
// fragmented the arena is.
class Arena
{
public:
    explicit Arena(size_t bytes) : size_(bytes)
    {
        insert_free(0, bytes);
    }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes   = round(bytes);
        auto it = by_size_.lower_bound({bytes, size_t(0)});
        if(it == by_size_.end())
            return false;
        const size_t boff = it->second, bsize = it->first;
        erase_free(boff, bsize);
        if(bsize > bytes)
            insert_free(boff + bytes, bsize - bytes);
        *off = boff;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = round(bytes);
        used_ -= bytes;
        auto nx = free_.lower_bound(off);
        if(nx != free_.end() && off + bytes == nx->first)
        {
            bytes += nx->second;
            erase_free(nx->first, nx->second);
        }
        auto pv = free_.lower_bound(off);
        if(pv != free_.begin())
        {
            --pv;
            if(pv->first + pv->second == off)
            {
                off = pv->first;
                bytes += pv->second;
                erase_free(pv->first, pv->second);
            }
        }
        insert_free(off, bytes);
    }
    size_t peak() const
    {
        return peak_;
    }
    size_t size() const
    {
        return size_;
    }

private:
    static size_t round(size_t b)
    {
        return (b + 255) / 256 * 256;
    }
    void insert_free(size_t off, size_t bytes)
    {
        free_[off] = bytes;
        by_size_.emplace(bytes, off);
    }
    void erase_free(size_t off, size_t bytes)
    {
        free_.erase(off);
        by_size_.erase({bytes, off});
    }
    size_t                              size_;
    size_t                              used_ = 0;
    size_t                              peak_ = 0;
    std::map<size_t, size_t>            free_;    // offset -> bytes
    std::set<std::pair<size_t, size_t>> by_size_; // (bytes, offset)
};

struct Factor
{
    size_t     off       = 0;
    size_t     bytes     = 0;
    int        m         = 0;
    int        prec      = 0;
    int        users     = 0;       // admitted, not yet retired requests (pins it)
    uint64_t   last      = 0;       // LRU stamp
    bool       generated = false;   // synthesis launched
    int        gen_stream = -1;     // stream the synthesis ran on
    hipEvent_t ready     = nullptr; // recorded after the synthesis
};

struct Launch
{
    hipEvent_t           done = nullptr;
    std::vector<int64_t> reqs;
    // vbatched argument block (device arena) and its pinned host staging
    size_t args_off   = 0;
    size_t args_bytes = 0;
    void*  host_args  = nullptr;
    size_t host_bytes = 0;
};

size_t elt(int prec)
{
    return prec == 0 ? sizeof(double) : sizeof(float);
}
} // namespace

struct rocblas_amd_server_
{
    int                         device       = 0;
    char*                       base         = nullptr;
    Arena*                      arena        = nullptr;
    std::vector<hipStream_t>    streams;
    std::vector<rocblas_handle> handles;
    int                         next_stream  = 0;
    int                         max_inflight = 256;
    int                         small_k      = 512;
    std::unordered_map<int64_t, Factor> factors;
    // resident factors no admitted request uses, least recently used first:
    // (last, factor id); the eviction candidates
    std::set<std::pair<uint64_t, int64_t>> idle;
    uint64_t                    stamp = 0;
    // the current run's requests (copied) and per-request device state
    std::vector<rocblas_amd_request> reqv;
    const rocblas_amd_request*       reqs = nullptr;
    int64_t                          nreq = 0;
    std::vector<size_t>              b_off;
    std::vector<char>                b_live;
    std::vector<char>                kept;
    rocblas_amd_serve_stats          stats{};
    // recycled pinned staging buffers (bytes, pointer) and events
    std::vector<std::pair<size_t, void*>> pinned_free;
    std::vector<hipEvent_t>               events_free;
};

namespace
{
using Server = rocblas_amd_server_;

void* dev_ptr(Server* s, size_t off)
{
    return s->base + off;
}

// A factor is pinned from admission to retirement of each request using it;
// unpinned factors are the eviction candidates.
void pin(Server* s, int64_t id, Factor& f)
{
    if(f.users++ == 0)
        s->idle.erase({f.last, id});
}

void unpin(Server* s, int64_t id, Factor& f, bool touch)
{
    if(--f.users == 0)
    {
        if(touch)
            f.last = ++s->stamp;
        s->idle.emplace(f.last, id);
    }
}

bool evict_for(Server* s, size_t bytes, size_t* off)
{
    while(!s->arena->alloc(bytes, off))
    {
        // least recently used factor with no in-flight user
        if(s->idle.empty())
            return false;
        auto victim = s->factors.find(s->idle.begin()->second);
        s->idle.erase(s->idle.begin());
        s->arena->release(victim->second.off, victim->second.bytes);
        if(victim->second.ready)
            (void)hipEventDestroy(victim->second.ready);
        s->factors.erase(victim);
        s->stats.evictions++;
    }
    return true;
}

rocblas_status synth_factor(rocblas_handle h, const rocblas_amd_request& r, void* A)
{
    if(r.precision == 0)
        return rocblas_amd_dsynth_tri(h, rocblas_fill_lower, rocblas_amd_synth_uniform, r.m,
                                      r.factor_seed, 1.0, 0.0, static_cast<double*>(A), r.m);
    return rocblas_amd_ssynth_tri(h, rocblas_fill_lower, rocblas_amd_synth_uniform, r.m,
                                  r.factor_seed, 1.0, 0.0, static_cast<float*>(A), r.m);
}

rocblas_status synth_rhs(rocblas_handle h, const rocblas_amd_request& r, void* B)
{
    if(r.precision == 0)
        return rocblas_amd_dsynth_rhs(h, r.m, r.n, r.rhs_seed, 2, static_cast<double*>(B), r.m);
    return rocblas_amd_ssynth_rhs(h, r.m, r.n, r.rhs_seed, 2, static_cast<float*>(B), r.m);
}

double seconds_since(clock_type::time_point t0)
{
    return std::chrono::duration<double>(clock_type::now() - t0).count();
}

bool pinned_acquire(Server* s, size_t bytes, void** p, size_t* got)
{
    for(size_t i = 0; i < s->pinned_free.size(); ++i)
        if(s->pinned_free[i].first >= bytes)
        {
            *got = s->pinned_free[i].first;
            *p   = s->pinned_free[i].second;
            s->pinned_free.erase(s->pinned_free.begin() + i);
            return true;
        }
    size_t sz = 4096;
    while(sz < bytes)
        sz *= 2;
    if(hipHostMalloc(p, sz, hipHostMallocDefault) != hipSuccess)
        return false;
    *got = sz;
    return true;
}

bool event_acquire(Server* s, hipEvent_t* e)
{
    if(!s->events_free.empty())
    {
        *e = s->events_free.back();
        s->events_free.pop_back();
        return true;
    }
    return hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
}

// Launch one group of admitted requests of one precision.  Batched groups of
// small solves go to stream 0, large solves rotate over the other streams, so
// a small request never queues behind a large solve on its stream.
rocblas_status launch_group(Server* s, const std::vector<int64_t>& group, bool batched, int prec,
                            std::vector<std::deque<Launch>>& inflight)
{
    const int ns = int(s->streams.size());
    int       si = 0;
    if(!batched && ns > 1)
    {
        si             = 1 + s->next_stream;
        s->next_stream = (s->next_stream + 1) % (ns - 1);
    }
    rocblas_handle h  = s->handles[si];
    Launch         L;
    L.reqs = group;
    rocblas_status st;
    // generate inputs (factor on a miss, right-hand side always) on this
    // stream; a factor generated on another stream is waited for
    for(int64_t id : group)
    {
        const rocblas_amd_request& r = s->reqs[id];
        Factor&                    f = s->factors[r.factor_id];
        if(!f.generated)
        {
            if((st = synth_factor(h, r, dev_ptr(s, f.off))) != rocblas_status_success)
                return st;
            if(!f.ready
               && hipEventCreateWithFlags(&f.ready, hipEventDisableTiming) != hipSuccess)
                return rocblas_status_internal_error;
            if(hipEventRecord(f.ready, s->streams[si]) != hipSuccess)
                return rocblas_status_internal_error;
            f.generated  = true;
            f.gen_stream = si;
        }
        else if(f.gen_stream != si
                && hipStreamWaitEvent(s->streams[si], f.ready, 0) != hipSuccess)
            return rocblas_status_internal_error;
        if((st = synth_rhs(h, r, dev_ptr(s, s->b_off[id]))) != rocblas_status_success)
            return st;
    }
    if(batched)
    {
        const int64_t nb   = int64_t(group.size());
        const size_t  ints = 4 * sizeof(int) * nb, ptrs = 2 * sizeof(void*) * nb;
        L.args_bytes       = ints + ptrs + 256;
        if(!evict_for(s, L.args_bytes, &L.args_off))
            return rocblas_status_memory_error;
        if(!pinned_acquire(s, L.args_bytes, &L.host_args, &L.host_bytes))
            return rocblas_status_memory_error;
        char*  hb  = static_cast<char*>(L.host_args);
        int*   mi  = reinterpret_cast<int*>(hb);
        int*   ni  = mi + nb;
        int*   la  = ni + nb;
        int*   lb  = la + nb;
        size_t pof = (ints + 15) / 16 * 16;
        void** pa  = reinterpret_cast<void**>(hb + pof);
        void** pb  = pa + nb;
        for(int64_t i = 0; i < nb; ++i)
        {
            const rocblas_amd_request& r = s->reqs[group[i]];
            mi[i] = r.m;
            ni[i] = r.n;
            la[i] = r.m;
            lb[i] = r.m;
            pa[i] = dev_ptr(s, s->factors[r.factor_id].off);
            pb[i] = dev_ptr(s, s->b_off[group[i]]);
        }
        char* db = static_cast<char*>(dev_ptr(s, L.args_off));
        if(hipMemcpyAsync(db, hb, pof + ptrs, hipMemcpyHostToDevice, s->streams[si])
           != hipSuccess)
            return rocblas_status_internal_error;
        const int* dm  = reinterpret_cast<const int*>(db);
        const int* dn  = dm + nb;
        const int* dla = dn + nb;
        const int* dlb = dla + nb;
        if(prec == 0)
        {
            const double one = 1.0;
            st = rocblas_amd_dtrsm_vbatched(
                h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                rocblas_diagonal_non_unit, dm, dn, &one,
                reinterpret_cast<const double* const*>(db + pof), dla,
                reinterpret_cast<double* const*>(db + pof + sizeof(void*) * nb), dlb, int(nb));
        }
        else
        {
            const float one = 1.0f;
            st = rocblas_amd_strsm_vbatched(
                h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                rocblas_diagonal_non_unit, dm, dn, &one,
                reinterpret_cast<const float* const*>(db + pof), dla,
                reinterpret_cast<float* const*>(db + pof + sizeof(void*) * nb), dlb, int(nb));
        }
        if(st != rocblas_status_success)
            return st;
        s->stats.batched_launches++;
    }
    else
    {
        for(int64_t id : group)
        {
            const rocblas_amd_request& r = s->reqs[id];
            void* A = dev_ptr(s, s->factors[r.factor_id].off);
            void* B = dev_ptr(s, s->b_off[id]);
            if(prec == 0)
            {
                const double one = 1.0;
                st = rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower,
                                   rocblas_operation_none, rocblas_diagonal_non_unit, r.m, r.n,
                                   &one, static_cast<const double*>(A), r.m,
                                   static_cast<double*>(B), r.m);
            }
            else
            {
                const float one = 1.0f;
                st = rocblas_strsm(h, rocblas_side_left, rocblas_fill_lower,
                                   rocblas_operation_none, rocblas_diagonal_non_unit, r.m, r.n,
                                   &one, static_cast<const float*>(A), r.m,
                                   static_cast<float*>(B), r.m);
            }
            if(st != rocblas_status_success)
                return st;
        }
    }
    if(!event_acquire(s, &L.done) || hipEventRecord(L.done, s->streams[si]) != hipSuccess)
        return rocblas_status_internal_error;
    s->stats.launches++;
    inflight[si].push_back(std::move(L));
    return rocblas_status_success;
}

void retire(Server* s, Launch& L, double now, double* latency)
{
    for(int64_t id : L.reqs)
    {
        const rocblas_amd_request& r = s->reqs[id];
        latency[id]                  = now - r.arrival;
        unpin(s, r.factor_id, s->factors[r.factor_id], true);
        if(r.keep)
            s->kept[id] = 1;
        else
        {
            s->arena->release(s->b_off[id], size_t(r.m) * r.n * elt(r.precision));
            s->b_live[id] = 0;
        }
        s->stats.completed++;
    }
    if(L.args_bytes)
        s->arena->release(L.args_off, L.args_bytes);
    if(L.host_args)
        s->pinned_free.emplace_back(L.host_bytes, L.host_args);
    s->events_free.push_back(L.done);
}

rocblas_status run(Server* s, const rocblas_amd_request* in, int64_t count, double* latency)
{
    // drop the previous run's kept results
    for(size_t i = 0; i < s->b_live.size(); ++i)
        if(s->b_live[i])
            s->arena->release(s->b_off[i],
                              size_t(s->reqs[i].m) * s->reqs[i].n * elt(s->reqs[i].precision));
    s->reqv.assign(in, in + count);
    const rocblas_amd_request* reqs = s->reqv.data();
    s->reqs                         = reqs;
    s->nreq                         = count;
    s->b_off.assign(count, 0);
    s->b_live.assign(count, 0);
    s->kept.assign(count, 0);
    s->stats = rocblas_amd_serve_stats{};
    for(int64_t i = 0; i < count; ++i)
    {
        latency[i] = -1.0;
        const auto& r = reqs[i];
        if(r.m <= 0 || r.n <= 0 || (r.precision != 0 && r.precision != 1) || r.arrival < 0.0
           || (i && r.arrival < reqs[i - 1].arrival))
            return rocblas_status_invalid_value;
    }

    std::deque<int64_t> pending;
    // launches in flight, per stream in launch order (a stream completes its
    // work in order, so only the oldest launch of each stream is polled)
    std::vector<std::deque<Launch>> inflight(s->streams.size());
    int64_t             next_arrival = 0, inflight_reqs = 0, done = 0;
    const auto          t0           = clock_type::now();
    rocblas_status      st           = rocblas_status_success;

    while(done < count)
    {
        double now = seconds_since(t0);
        while(next_arrival < count && reqs[next_arrival].arrival <= now)
            pending.push_back(next_arrival++);

        // completions
        bool progressed = false;
        for(auto& q : inflight)
            while(!q.empty())
            {
                const hipError_t e = hipEventQuery(q.front().done);
                if(e == hipErrorNotReady)
                    break;
                if(e != hipSuccess)
                    return rocblas_status_internal_error;
                const int64_t n = int64_t(q.front().reqs.size());
                retire(s, q.front(), seconds_since(t0), latency);
                q.pop_front();
                inflight_reqs -= n;
                done += n;
                progressed = true;
            }

        // admission (FIFO) under the arena budget and the in-flight cap
        std::vector<int64_t> small[2], large[2];
        while(!pending.empty() && inflight_reqs < s->max_inflight)
        {
            const int64_t              id = pending.front();
            const rocblas_amd_request& r  = reqs[id];
            const size_t               eb = elt(r.precision);
            const size_t               fbytes = size_t(r.m) * r.m * eb;
            const size_t               bbytes = size_t(r.m) * r.n * eb;
            if(fbytes + bbytes + 512 > s->arena->size())
            {
                // can never fit: reject
                pending.pop_front();
                latency[id] = -1.0;
                s->stats.rejected++;
                done++;
                continue;
            }
            auto fit = s->factors.find(r.factor_id);
            bool hit = fit != s->factors.end();
            if(hit && (fit->second.m != r.m || fit->second.prec != r.precision))
                return rocblas_status_invalid_value; // one factor id, one shape
            if(hit)
                pin(s, r.factor_id, fit->second); // pinned from admission to retirement
            // nothing in flight and still no room (fragmentation by kept
            // results): reject rather than wait for memory that never frees
            auto stuck = [&]() {
                if(inflight_reqs > 0)
                    return false;
                pending.pop_front();
                latency[id] = -1.0;
                s->stats.rejected++;
                done++;
                return true;
            };
            size_t boff = 0;
            if(!evict_for(s, bbytes, &boff))
            {
                if(hit)
                    unpin(s, r.factor_id, fit->second, false);
                if(stuck())
                    continue;
                break; // wait for completions
            }
            if(hit)
                s->stats.factor_hits++;
            else
            {
                size_t foff = 0;
                if(!evict_for(s, fbytes, &foff))
                {
                    s->arena->release(boff, bbytes);
                    if(stuck())
                        continue;
                    break;
                }
                Factor f;
                f.off   = foff;
                f.bytes = fbytes;
                f.m     = r.m;
                f.prec  = r.precision;
                f.users = 1;
                f.last  = ++s->stamp;
                s->factors[r.factor_id] = f;
                s->stats.factor_misses++;
            }
            s->b_off[id]  = boff;
            s->b_live[id] = 1;
            pending.pop_front();
            inflight_reqs++;
            (r.m <= s->small_k ? small : large)[r.precision].push_back(id);
        }
        for(int p = 0; p < 2; ++p)
        {
            if(!small[p].empty())
            {
                if((st = launch_group(s, small[p], true, p, inflight)) != rocblas_status_success)
                    return st;
                progressed = true;
            }
            for(int64_t id : large[p])
            {
                if((st = launch_group(s, std::vector<int64_t>{id}, false, p, inflight))
                   != rocblas_status_success)
                    return st;
                progressed = true;
            }
        }
        if(!progressed)
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    s->stats.wall_seconds = seconds_since(t0);

    // latency percentiles over the completed requests
    std::vector<double> lat;
    lat.reserve(count);
    double sum = 0.0;
    for(int64_t i = 0; i < count; ++i)
        if(latency[i] >= 0.0)
        {
            lat.push_back(latency[i]);
            sum += latency[i];
        }
    if(!lat.empty())
    {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) {
            const size_t i = size_t(q * double(lat.size() - 1) + 0.5);
            return lat[std::min(i, lat.size() - 1)];
        };
        s->stats.p50_latency  = pct(0.50);
        s->stats.p99_latency  = pct(0.99);
        s->stats.max_latency  = lat.back();
        s->stats.mean_latency = sum / double(lat.size());
    }
    s->stats.peak_arena_bytes = double(s->arena->peak());
    return rocblas_status_success;
}
} // namespace

extern "C" {
rocblas_status rocblas_amd_server_create(rocblas_amd_server* server,
                                         size_t              hbm_budget_bytes,
                                         rocblas_int         streams,
                                         rocblas_int         max_inflight,
                                         rocblas_int         small_k)
try
{
    if(!server)
        return rocblas_status_invalid_pointer;
    *server = nullptr;
    if(hbm_budget_bytes < 4096 || streams < 1 || streams > 16 || max_inflight < 1
       || small_k < 0)
        return rocblas_status_invalid_size;
    Server* s = new Server;
    s->max_inflight = max_inflight;
    s->small_k      = small_k;
    if(hipGetDevice(&s->device) != hipSuccess
       || hipMalloc(reinterpret_cast<void**>(&s->base), hbm_budget_bytes) != hipSuccess)
    {
        (void)hipGetLastError();
        delete s;
        return rocblas_status_memory_error;
    }
    s->arena = new Arena(hbm_budget_bytes);
    for(int i = 0; i < streams; ++i)
    {
        hipStream_t    st = nullptr;
        rocblas_handle h  = nullptr;
        if(hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess
           || rocblas_create_handle(&h) != rocblas_status_success
           || rocblas_set_stream(h, st) != rocblas_status_success)
        {
            rocblas_amd_server_destroy(s);
            return rocblas_status_internal_error;
        }
        s->streams.push_back(st);
        s->handles.push_back(h);
    }
    *server = s;
    return rocblas_status_success;
}
catch(...)
{
    return rocblas_status_memory_error;
}

rocblas_status rocblas_amd_server_destroy(rocblas_amd_server server)
{
    if(!server)
        return rocblas_status_invalid_pointer;
    for(hipStream_t st : server->streams)
        (void)hipStreamSynchronize(st);
    for(auto& f : server->factors)
        if(f.second.ready)
            (void)hipEventDestroy(f.second.ready);
    for(auto& pb : server->pinned_free)
        (void)hipHostFree(pb.second);
    for(hipEvent_t e : server->events_free)
        (void)hipEventDestroy(e);
    for(rocblas_handle h : server->handles)
        rocblas_destroy_handle(h);
    for(hipStream_t st : server->streams)
        (void)hipStreamDestroy(st);
    if(server->base)
        (void)hipFree(server->base);
    delete server->arena;
    delete server;
    return rocblas_status_success;
}

rocblas_status rocblas_amd_server_run(rocblas_amd_server         server,
                                      const rocblas_amd_request* requests,
                                      int64_t                    count,
                                      double*                    latency,
                                      rocblas_amd_serve_stats*   stats)
try
{
    if(!server)
        return rocblas_status_invalid_pointer;
    if(count < 0)
        return rocblas_status_invalid_size;
    if(count && (!requests || !latency))
        return rocblas_status_invalid_pointer;
    rbamd::DeviceGuard dg(server->device);
    rocblas_status     st = run(server, requests, count, latency);
    for(hipStream_t s : server->streams)
        (void)hipStreamSynchronize(s);
    if(stats)
        *stats = server->stats;
    return st;
}
catch(const std::bad_alloc&)
{
    return rocblas_status_memory_error;
}
catch(...)
{
    return rocblas_status_internal_error;
}

rocblas_status rocblas_amd_server_result(rocblas_amd_server server, int64_t index, void* X,
                                         rocblas_int ldx)
{
    if(!server || !X)
        return rocblas_status_invalid_pointer;
    if(index < 0 || index >= server->nreq || !server->kept[index] || !server->b_live[index])
        return rocblas_status_invalid_value;
    const rocblas_amd_request& r = server->reqs[index];
    if(ldx < r.m)
        return rocblas_status_invalid_size;
    const size_t e = elt(r.precision);
    if(hipMemcpy2D(X, size_t(ldx) * e, server->base + server->b_off[index], size_t(r.m) * e,
                   size_t(r.m) * e, r.n, hipMemcpyDeviceToHost)
       != hipSuccess)
        return rocblas_status_internal_error;
    return rocblas_status_success;
}
}
