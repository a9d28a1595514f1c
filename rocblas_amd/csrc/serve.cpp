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
// Each request has its own side, uplo, transA, diag and alpha; the batched
// launches take a case and an alpha per instance.  The factor comes from
// rocblas_amd_{s,d}synth_tri (uniform) under the factor's seed and the
// right-hand side from rocblas_amd_{s,d}synth_rhs under the request's seed,
// so a request's result does not depend on what else was in flight: every
// kernel it goes through treats instances independently.
//
// Verification (verify != 0): the right-hand side is generated twice, into B
// (solved in place) and B0, and after the solve of a group the batched
// backward-error kernels (residual.hip) compute every request's
// ||alpha B0 - op(A) X||_1 / (||op(A)||_1 ||X||_1 + |alpha| ||B0||_1) in FP64 on
// the same stream; the values are copied to pinned host memory before the
// group's completion event, so they are read when the request retires.
#include "internal.hpp"
#include "residual.hpp"
#include "vbatched.hpp"

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
using namespace rbamd;
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
    int        uplo      = 0;
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
    // argument block (device arena) and its pinned host staging
    size_t  args_off   = 0;
    size_t  args_bytes = 0;
    void*   host_args  = nullptr;
    size_t  host_bytes = 0;
    double* host_eta   = nullptr; // backward errors, copied back (verify)
};

size_t elt(int prec)
{
    return prec == 0 ? sizeof(double) : sizeof(float);
}

size_t round256(size_t b)
{
    return (b + 255) / 256 * 256;
}

// rows x cols of B for a request (m: triangle order, n: vectors)
int b_rows(const rocblas_amd_request& r)
{
    return r.side == rocblas_side_left ? r.m : r.n;
}
int b_cols(const rocblas_amd_request& r)
{
    return r.side == rocblas_side_left ? r.n : r.m;
}
size_t b_bytes(const rocblas_amd_request& r)
{
    return round256(size_t(r.m) * r.n * elt(r.precision));
}

// Sections of one launch's argument block, in bytes from its start; the
// layout is the same in the pinned staging and in the device copy.
struct ArgLayout
{
    size_t vm, vn, kk, ld, fl, nv, alpha, pa, pb, pb0, acc_off, items, h2d_bytes;
    size_t acc, eta, total;
    explicit ArgLayout(int64_t nb, int64_t nitems, int64_t acc_doubles, size_t et)
    {
        auto take = [&](size_t& at, size_t bytes) {
            at = total;
            total += (bytes + 15) / 16 * 16;
        };
        total = 0;
        take(vm, 4 * nb);
        take(vn, 4 * nb);
        take(kk, 4 * nb);
        take(ld, 4 * nb);
        take(fl, 4 * nb);
        take(nv, 4 * nb);
        take(alpha, et * nb);
        take(pa, sizeof(void*) * nb);
        take(pb, sizeof(void*) * nb);
        take(pb0, sizeof(void*) * nb);
        take(acc_off, 8 * nb);
        take(items, 12 * nitems);
        h2d_bytes = total;
        take(acc, 8 * acc_doubles);
        take(eta, 8 * nb);
    }
};
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
    bool                        verify       = true;
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
    double*                          eta_out = nullptr; // caller's array (may be null)
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
    const rocblas_fill uplo = rocblas_fill(r.uplo);
    if(r.precision == 0)
        return rocblas_amd_dsynth_tri(h, uplo, rocblas_amd_synth_uniform, r.m, r.factor_seed, 1.0,
                                      0.0, static_cast<double*>(A), r.m);
    return rocblas_amd_ssynth_tri(h, uplo, rocblas_amd_synth_uniform, r.m, r.factor_seed, 1.0,
                                  0.0, static_cast<float*>(A), r.m);
}

rocblas_status synth_rhs(rocblas_handle h, const rocblas_amd_request& r, void* B)
{
    const int rows = b_rows(r), cols = b_cols(r);
    if(r.precision == 0)
        return rocblas_amd_dsynth_rhs(h, rows, cols, r.rhs_seed, 2, static_cast<double*>(B),
                                      rows);
    return rocblas_amd_ssynth_rhs(h, rows, cols, r.rhs_seed, 2, static_cast<float*>(B), rows);
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

template <typename T>
rocblas_status solve_one(rocblas_handle h, const rocblas_amd_request& r, const void* A, void* B)
{
    const T alpha = T(r.alpha);
    if constexpr(sizeof(T) == sizeof(double))
        return rocblas_dtrsm(h, rocblas_side(r.side), rocblas_fill(r.uplo),
                             rocblas_operation(r.transA), rocblas_diagonal(r.diag), b_rows(r),
                             b_cols(r), &alpha, static_cast<const double*>(A), r.m,
                             static_cast<double*>(B), b_rows(r));
    else
        return rocblas_strsm(h, rocblas_side(r.side), rocblas_fill(r.uplo),
                             rocblas_operation(r.transA), rocblas_diagonal(r.diag), b_rows(r),
                             b_cols(r), &alpha, static_cast<const float*>(A), r.m,
                             static_cast<float*>(B), b_rows(r));
}

// Solves (and, with verify, measures) one group of one precision: the
// argument block is staged in pinned memory and copied to the arena on the
// group's stream.
template <typename T>
rocblas_status solve_group(Server* s, Launch& L, int si, bool batched)
{
    rocblas_handle     h    = s->handles[si];
    hipStream_t        strm = s->streams[si];
    const auto&        grp  = L.reqs;
    const int64_t      nb   = int64_t(grp.size());
    int64_t            nitems = 0, acc_doubles = 0;
    if(s->verify)
        for(int64_t id : grp)
        {
            const rocblas_amd_request& r = s->reqs[id];
            nitems += eta_items(r.m, r.n);
            acc_doubles += eta_acc_doubles(r.m, r.n, r.side == rocblas_side_left);
        }
    const ArgLayout lay(nb, nitems, acc_doubles, sizeof(T));
    L.args_bytes = lay.total;
    if(!evict_for(s, L.args_bytes, &L.args_off))
        return rocblas_status_memory_error;
    if(!pinned_acquire(s, lay.total, &L.host_args, &L.host_bytes))
        return rocblas_status_memory_error;
    char* hb = static_cast<char*>(L.host_args);
    char* db = static_cast<char*>(dev_ptr(s, L.args_off));
    auto  hi = [&](size_t at) { return reinterpret_cast<int*>(hb + at); };
    int*      vm   = hi(lay.vm);
    int*      vn   = hi(lay.vn);
    int*      kk   = hi(lay.kk);
    int*      ld   = hi(lay.ld);
    int*      fl   = hi(lay.fl);
    int*      nv   = hi(lay.nv);
    T*        al   = reinterpret_cast<T*>(hb + lay.alpha);
    void**    pa   = reinterpret_cast<void**>(hb + lay.pa);
    void**    pb   = reinterpret_cast<void**>(hb + lay.pb);
    void**    pb0  = reinterpret_cast<void**>(hb + lay.pb0);
    int64_t*  aoff = reinterpret_cast<int64_t*>(hb + lay.acc_off);
    int*      it   = hi(lay.items);
    int64_t   acc_at = 0;
    for(int64_t i = 0; i < nb; ++i)
    {
        const rocblas_amd_request& r = s->reqs[grp[i]];
        vm[i]  = b_rows(r);
        vn[i]  = b_cols(r);
        kk[i]  = r.m;
        ld[i]  = b_rows(r);
        fl[i]  = trsm_case_flags(rocblas_side(r.side), rocblas_fill(r.uplo),
                                rocblas_operation(r.transA), rocblas_diagonal(r.diag));
        nv[i]  = r.n;
        al[i]  = T(r.alpha);
        pa[i]  = dev_ptr(s, s->factors[r.factor_id].off);
        pb[i]  = dev_ptr(s, s->b_off[grp[i]]);
        pb0[i] = dev_ptr(s, s->b_off[grp[i]] + b_bytes(r));
        aoff[i] = acc_at;
        if(s->verify)
        {
            acc_at += eta_acc_doubles(r.m, r.n, r.side == rocblas_side_left);
            eta_fill_items(int(i), r.m, r.n, it);
            it += 3 * eta_items(r.m, r.n);
        }
    }
    if(hipMemcpyAsync(db, hb, lay.h2d_bytes, hipMemcpyHostToDevice, strm) != hipSuccess)
        return rocblas_status_internal_error;
    rocblas_status st;
    auto di = [&](size_t at) { return reinterpret_cast<const int*>(db + at); };
    if(batched)
    {
        st = vbatched_cases<T>(h, di(lay.vm), di(lay.vn), di(lay.kk), di(lay.ld),
                               reinterpret_cast<const T* const*>(db + lay.pa),
                               reinterpret_cast<T* const*>(db + lay.pb), di(lay.fl),
                               reinterpret_cast<const T*>(db + lay.alpha), int(nb));
        if(st != rocblas_status_success)
            return st;
        s->stats.batched_launches++;
    }
    else
        for(int64_t i = 0; i < nb; ++i)
            if((st = solve_one<T>(h, s->reqs[grp[i]], pa[i], pb[i])) != rocblas_status_success)
                return st;
    if(s->verify)
    {
        EtaArgs<T> e{};
        e.K         = di(lay.kk);
        e.nvec      = di(lay.nv);
        e.flags     = di(lay.fl);
        e.lda       = di(lay.kk);
        e.ldx       = di(lay.ld);
        e.ldb       = di(lay.ld);
        e.A         = reinterpret_cast<const T* const*>(db + lay.pa);
        e.X         = reinterpret_cast<const T* const*>(db + lay.pb);
        e.B0        = reinterpret_cast<const T* const*>(db + lay.pb0);
        e.alpha     = reinterpret_cast<const T*>(db + lay.alpha);
        e.acc_off   = reinterpret_cast<const int64_t*>(db + lay.acc_off);
        e.acc       = reinterpret_cast<double*>(db + lay.acc);
        e.acc_total = acc_doubles;
        e.items     = di(lay.items);
        e.nitems    = int(nitems);
        e.batch     = int(nb);
        e.eta       = reinterpret_cast<double*>(db + lay.eta);
        if((st = eta_launch<T>(strm, e)) != rocblas_status_success)
            return st;
        L.host_eta = reinterpret_cast<double*>(hb + lay.eta);
        if(hipMemcpyAsync(L.host_eta, db + lay.eta, 8 * nb, hipMemcpyDeviceToHost, strm)
           != hipSuccess)
            return rocblas_status_internal_error;
    }
    return rocblas_status_success;
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
    // generate inputs (factor on a miss, right-hand side always, twice with
    // verification) on this stream; a factor generated on another stream is
    // waited for
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
        if(s->verify
           && (st = synth_rhs(h, r, dev_ptr(s, s->b_off[id] + b_bytes(r))))
                  != rocblas_status_success)
            return st;
    }
    st = prec == 0 ? solve_group<double>(s, L, si, batched) : solve_group<float>(s, L, si, batched);
    if(st != rocblas_status_success)
        return st;
    if(!event_acquire(s, &L.done) || hipEventRecord(L.done, s->streams[si]) != hipSuccess)
        return rocblas_status_internal_error;
    s->stats.launches++;
    inflight[si].push_back(std::move(L));
    return rocblas_status_success;
}

size_t b_alloc(const Server* s, const rocblas_amd_request& r)
{
    return (s->verify ? 2 : 1) * b_bytes(r);
}

void retire(Server* s, Launch& L, double now, double* latency)
{
    for(size_t j = 0; j < L.reqs.size(); ++j)
    {
        const int64_t              id = L.reqs[j];
        const rocblas_amd_request& r  = s->reqs[id];
        latency[id]                   = now - r.arrival;
        unpin(s, r.factor_id, s->factors[r.factor_id], true);
        if(r.keep)
            s->kept[id] = 1;
        else
        {
            s->arena->release(s->b_off[id], b_alloc(s, r));
            s->b_live[id] = 0;
        }
        if(L.host_eta)
        {
            const double e = L.host_eta[j];
            if(s->eta_out)
                s->eta_out[id] = e;
            s->stats.checked++;
            double& worst = r.precision == 0 ? s->stats.max_backward_error_fp64
                                             : s->stats.max_backward_error_fp32;
            worst = std::max(worst, e);
            if(!(e <= (r.precision == 0 ? 1e-12 : 1e-5)))
                s->stats.over_tolerance++;
        }
        s->stats.completed++;
    }
    if(L.args_bytes)
        s->arena->release(L.args_off, L.args_bytes);
    if(L.host_args)
        s->pinned_free.emplace_back(L.host_bytes, L.host_args);
    s->events_free.push_back(L.done);
}

rocblas_status run(Server* s, const rocblas_amd_request* in, int64_t count, double* latency,
                   double* eta_out)
{
    // drop the previous run's kept results
    for(size_t i = 0; i < s->b_live.size(); ++i)
        if(s->b_live[i])
            s->arena->release(s->b_off[i], b_alloc(s, s->reqs[i]));
    s->b_live.clear();
    s->reqv.assign(in, in + count);
    s->nreq = 0;
    // 0 selects the default case: left, lower, no transpose, non-unit
    for(auto& r : s->reqv)
    {
        r.side   = r.side ? r.side : rocblas_side_left;
        r.uplo   = r.uplo ? r.uplo : rocblas_fill_lower;
        r.transA = r.transA ? r.transA : rocblas_operation_none;
        r.diag   = r.diag ? r.diag : rocblas_diagonal_non_unit;
    }
    const rocblas_amd_request* reqs = s->reqv.data();
    s->reqs                         = reqs;
    s->eta_out                      = eta_out;
    s->stats = rocblas_amd_serve_stats{};
    for(int64_t i = 0; i < count; ++i)
    {
        latency[i] = -1.0;
        if(eta_out)
            eta_out[i] = -1.0;
        const auto& r = reqs[i];
        if(r.m <= 0 || r.n <= 0 || (r.precision != 0 && r.precision != 1) || r.arrival < 0.0
           || (i && r.arrival < reqs[i - 1].arrival)
           || (r.side != rocblas_side_left && r.side != rocblas_side_right)
           || (r.uplo != rocblas_fill_lower && r.uplo != rocblas_fill_upper)
           || (r.transA != rocblas_operation_none && r.transA != rocblas_operation_transpose
               && r.transA != rocblas_operation_conjugate_transpose)
           || (r.diag != rocblas_diagonal_non_unit && r.diag != rocblas_diagonal_unit))
            return rocblas_status_invalid_value;
    }
    s->nreq = count;
    s->b_off.assign(count, 0);
    s->b_live.assign(count, 0);
    s->kept.assign(count, 0);

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
            const size_t               bbytes = b_alloc(s, r);
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
            if(hit
               && (fit->second.m != r.m || fit->second.prec != r.precision
                   || fit->second.uplo != r.uplo))
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
                f.uplo  = r.uplo;
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
                                         rocblas_int         small_k,
                                         rocblas_int         verify)
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
    s->verify       = verify != 0;
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
                                      double*                    backward_error,
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
    rocblas_status     st = run(server, requests, count, latency, backward_error);
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
    const int                  rows = b_rows(r), cols = b_cols(r);
    if(ldx < rows)
        return rocblas_status_invalid_size;
    const size_t e = elt(r.precision);
    if(hipMemcpy2D(X, size_t(ldx) * e, server->base + server->b_off[index], size_t(rows) * e,
                   size_t(rows) * e, cols, hipMemcpyDeviceToHost)
       != hipSuccess)
        return rocblas_status_internal_error;
    return rocblas_status_success;
}
}
