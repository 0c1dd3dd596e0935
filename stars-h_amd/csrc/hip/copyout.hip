// Copy-back of the factors of one API call, overlapped with the compression
// of later batches.
//
// A call compresses its tiles in several batches.  When a batch is done,
// the factors of its finished tiles are packed on the device into the
// call's arena (rsdd_out_append) and queued for the host.  A dispatcher
// thread moves each queued range device-to-host on a copy stream, through a
// ring of pinned slots, as soon as the pack kernel has run; worker threads
// copy every slot that has arrived into the caller's host buffers and give
// the slot back.  Meanwhile the calling thread goes on launching the next
// batch on the library stream.  rsdd_out_end waits for the last copies.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>
#include "hip_common.h"
#include "rsdd.h"

namespace {

__global__ __launch_bounds__(256) void arena_pack_kernel(const double *src,
        double *dst, const int64_t *seg, int nseg)
{
    // seg[3 s .. 3 s + 2] = source offset, arena offset, length
    const int s = blockIdx.y;
    if(s >= nseg)
        return;
    const int64_t so = seg[3*s], po = seg[3*s+1], len = seg[3*s+2];
    for(int64_t e = (int64_t)blockIdx.x*blockDim.x+threadIdx.x; e < len;
            e += (int64_t)gridDim.x*blockDim.x)
        dst[po+e] = src[so+e];
}

constexpr int NSLOT_MAX = 512;
constexpr int64_t RING = (int64_t)1 << 26;   // doubles of the ring (512 MiB)
// slot size (doubles) and count; starsh_hip_set_option("copy_slot_mb")
int64_t SLOT = (int64_t)1 << 22;             // 32 MiB
int NSLOT = (int)(RING/SLOT);
// Transfers on the library stream (1, default) or on a copy stream of
// their own (0, starsh_hip_set_option("copy_stream")).  On MI355X the
// library stream moves the C2 factors device-to-host at about 55 GB/s, the
// second stream at about 30 GB/s, which costs more than its overlap with
// a later batch saves; on the library stream a transfer runs after the
// kernels queued before it.
int use_lib_stream = 1;

struct Seg
{
    int64_t aoff, len;
    double *dst;
};

struct Job
{
    const double *src = nullptr;    // device source of the job's range
    int64_t lo = 0, hi = 0;        // range (arena offsets, or 0..len)
    hipEvent_t ready = nullptr;     // recorded after the pack kernel
    std::vector<Seg> segs;          // increasing aoff
    std::vector<char> touched;      // per TOUCH range: pages in (under mu)
};

// Work of the host threads: touch the destination pages of a range of a
// job (slot < 0; the first touch maps and zeroes a page, done while the
// transfer is still on its way), or copy a range of a slot that has
// arrived into the destinations.  Ranges are job-relative arena offsets.
struct Task
{
    int slot;
    int64_t off, len;
    int64_t slot_off;              // start of the slot's range (copy)
    const Job *job;
};

constexpr int NSUB = 4;            // copy tasks per slot
constexpr int64_t TOUCH = (int64_t)1 << 19;   // doubles per touch task

struct Engine
{
    // device arena and its fill level (doubles)
    double *arena = nullptr;
    int64_t arena_cap = 0, used = 0;
    // pinned segment lists of the pack kernels (int64 triples)
    int64_t *hseg = nullptr, *dseg = nullptr;
    int64_t seg_cap = 0, seg_used = 0;
    // pinned ring and its events
    double *ring = nullptr;
    hipEvent_t slot_ev[NSLOT_MAX] = {};
    bool slot_free[NSLOT_MAX];
    hipStream_t cs = nullptr;
    // queues
    std::mutex mu;
    std::condition_variable cv_job, cv_piece, cv_slot, cv_touch;
    std::deque<std::unique_ptr<Job>> jobs;   // all jobs of the call
    size_t next_job = 0;
    std::deque<Task> tasks;
    int slot_left[NSLOT_MAX];        // copy tasks of a slot not yet done
    bool ending = false, dispatch_done = false;
    int error = STARSH_SUCCESS;
    std::thread dispatcher;
    std::vector<std::thread> workers;
    bool running = false;
    // begin of the call, first queued batch, last copy done
    std::chrono::steady_clock::time_point t_first, t_job0, t_last;
    bool any = false;
};

Engine E;

// Optional timeline of the copy-back (STARSH_HIP_COPY_TRACE=file): one line
// per event, "kind slot offset length microseconds-since-begin".
struct TraceRec
{
    char kind;
    int slot;
    int64_t off, len;
    double us;
};
std::mutex trace_mu;
std::vector<TraceRec> trace;
bool trace_on = false;

void trace_add(char kind, int slot, int64_t off, int64_t len)
{
    if(!trace_on)
        return;
    const double us = std::chrono::duration<double, std::micro>(
            std::chrono::steady_clock::now()-E.t_first).count();
    std::lock_guard<std::mutex> g(trace_mu);
    trace.push_back(TraceRec{kind, slot, off, len, us});
}

void set_error(int e)
{
    std::lock_guard<std::mutex> g(E.mu);
    if(E.error == STARSH_SUCCESS)
        E.error = e;
}

// Applies f(dst, job-relative offset x, length) to the parts of the job's
// destination segments inside [a, b).
template <typename F>
void for_segments(const Job &j, int64_t a, int64_t b, F f)
{
    // first segment ending after a
    auto it = std::upper_bound(j.segs.begin(), j.segs.end(), a,
            [](int64_t v, const Seg &s) { return v < s.aoff; });
    size_t q = it == j.segs.begin() ? 0 : (size_t)(it-j.segs.begin())-1;
    for(; q < j.segs.size() && j.segs[q].aoff < b; q++)
    {
        const Seg &s = j.segs[q];
        const int64_t x = std::max(a, s.aoff), y = std::min(b, s.aoff+s.len);
        if(x < y)
            f(s.dst+(x-s.aoff), x, y-x);
    }
}

void run_task(const Task &tk)
{
    if(tk.slot < 0)
    {
        // the first write to a page maps and zeroes it; every copy task of
        // the job waits until all its touch tasks are done, so these
        // writes never land on copied data
        for_segments(*tk.job, tk.off, tk.off+tk.len,
                [](double *d, int64_t, int64_t len) {
                    volatile char *p = (volatile char *)d;
                    const int64_t nbytes = (int64_t)sizeof(double)*len;
                    for(int64_t o = 0; o < nbytes; o += 4096)
                        p[o] = 0;
                });
        return;
    }
    const double *h = E.ring+(int64_t)tk.slot*SLOT;
    const int64_t s0 = tk.slot_off;
    for_segments(*tk.job, tk.off, tk.off+tk.len,
            [&](double *d, int64_t x, int64_t len) {
                memcpy(d, h+(x-s0), sizeof(double)*(size_t)len);
            });
}

void worker_loop()
{
    for(;;)
    {
        Task tk;
        {
            std::unique_lock<std::mutex> g(E.mu);
            E.cv_piece.wait(g, [] { return !E.tasks.empty() ||
                    E.dispatch_done; });
            if(E.tasks.empty())
                return;
            tk = E.tasks.front();
            E.tasks.pop_front();
        }
        if(tk.slot >= 0)
        {
            // the pages of this range must be touched first (the touch
            // tasks of a job were queued before any of its copy tasks)
            const int64_t t0 = tk.off/TOUCH, t1 = (tk.off+tk.len-1)/TOUCH;
            std::unique_lock<std::mutex> g(E.mu);
            E.cv_touch.wait(g, [&] {
                for(int64_t q = t0; q <= t1; q++)
                    if(!tk.job->touched[q])
                        return false;
                return true;
            });
        }
        if(tk.slot >= 0)
            trace_add('w', tk.slot, tk.off, tk.len);
        if(tk.slot >= 0 && hipEventSynchronize(E.slot_ev[tk.slot]) !=
                hipSuccess)
            set_error(STARSH_UNKNOWN_ERROR);
        else
        {
            trace_add(tk.slot >= 0 ? 'a' : 't', tk.slot, tk.off, tk.len);
            run_task(tk);
            trace_add(tk.slot >= 0 ? 'c' : 'u', tk.slot, tk.off, tk.len);
        }
        if(tk.slot < 0)
        {
            {
                std::lock_guard<std::mutex> g(E.mu);
                const_cast<Job *>(tk.job)->touched[tk.off/TOUCH] = 1;
            }
            E.cv_touch.notify_all();
        }
        if(tk.slot >= 0)
        {
            bool freed = false;
            {
                std::lock_guard<std::mutex> g(E.mu);
                if(--E.slot_left[tk.slot] == 0)
                {
                    E.slot_free[tk.slot] = true;
                    freed = true;
                }
                E.t_last = std::chrono::steady_clock::now();
            }
            if(freed)
                E.cv_slot.notify_one();
        }
    }
}

void dispatcher_loop()
{
    for(;;)
    {
        Job *j;
        {
            std::unique_lock<std::mutex> g(E.mu);
            E.cv_job.wait(g, [] { return E.next_job < E.jobs.size() ||
                    E.ending; });
            if(E.next_job >= E.jobs.size())
                break;
            j = E.jobs[E.next_job++].get();
        }
        trace_add('D', -1, j->lo, j->hi-j->lo);
        hipStream_t cs = use_lib_stream ? starsh_hip_stream() : E.cs;
        if(!use_lib_stream && hipStreamWaitEvent(cs, j->ready, 0) !=
                hipSuccess)
        {
            set_error(STARSH_UNKNOWN_ERROR);
            continue;
        }
        trace_add('W', -1, j->lo, j->hi-j->lo);
        for(int64_t off = j->lo; off < j->hi; off += SLOT)
        {
            const int64_t len = std::min(SLOT, j->hi-off);
            int slot = -1;
            {
                std::unique_lock<std::mutex> g(E.mu);
                E.cv_slot.wait(g, [&] {
                    for(int s = 0; s < NSLOT; s++)
                        if(E.slot_free[s])
                        {
                            slot = s;
                            return true;
                        }
                    return false;
                });
                E.slot_free[slot] = false;
            }
            trace_add('d', slot, off, len);
            if(hipMemcpyAsync(E.ring+(int64_t)slot*SLOT, j->src+(off-j->lo),
                        sizeof(double)*(size_t)len, hipMemcpyDeviceToHost,
                        cs) != hipSuccess ||
                    hipEventRecord(E.slot_ev[slot], cs) != hipSuccess)
            {
                set_error(STARSH_UNKNOWN_ERROR);
                std::lock_guard<std::mutex> g(E.mu);
                E.slot_free[slot] = true;
                continue;
            }
            {
                std::lock_guard<std::mutex> g(E.mu);
                const int64_t r0 = off-j->lo;
                E.slot_left[slot] = NSUB;
                for(int q = 0; q < NSUB; q++)
                {
                    const int64_t a = r0+len*q/NSUB, b = r0+len*(q+1)/NSUB;
                    E.tasks.push_back(Task{slot, a, b-a, r0, j});
                }
            }
            E.cv_piece.notify_all();
        }
    }
    {
        std::lock_guard<std::mutex> g(E.mu);
        E.dispatch_done = true;
    }
    E.cv_piece.notify_all();
}

int grow_device(double **p, int64_t *cap, int64_t want)
{
    if(want <= *cap && *p)
        return STARSH_SUCCESS;
    if(*p)
        (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const int64_t unit = (int64_t)1 << 23;    // 64 MiB of doubles
    const int64_t c = (want+unit-1)/unit*unit;
    if(hipMalloc((void **)p, sizeof(double)*(size_t)c) != hipSuccess)
    {
        *p = nullptr;
        fprintf(stderr, "STARSH HIP ERROR: hipMalloc of the factor arena "
                "(%lld doubles) failed\n", (long long)c);
        return STARSH_MALLOC_ERROR;
    }
    *cap = c;
    return STARSH_SUCCESS;
}

} // namespace

int rsdd_out_begin(int64_t max_doubles, int64_t max_segments)
{
    if(E.running)
        return STARSH_UNKNOWN_ERROR;
    int info = grow_device(&E.arena, &E.arena_cap, max_doubles > 0 ?
            max_doubles : 1);
    if(info != STARSH_SUCCESS)
        return info;
    const int64_t ns = 3*(max_segments > 0 ? max_segments : 1);
    if(ns > E.seg_cap)
    {
        if(E.hseg)
            (void)hipHostFree(E.hseg);
        if(E.dseg)
            (void)hipFree(E.dseg);
        E.hseg = E.dseg = nullptr;
        E.seg_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&E.hseg, sizeof(int64_t)*ns,
                    hipHostMallocDefault));
        HIP_TRY(hipMalloc((void **)&E.dseg, sizeof(int64_t)*ns));
        E.seg_cap = ns;
    }
    if(E.ring == nullptr)
    {
        HIP_TRY(hipHostMalloc((void **)&E.ring,
                    sizeof(double)*(size_t)RING,
                    hipHostMallocDefault));
        for(int s = 0; s < NSLOT_MAX; s++)
            // blocking waits: the worker threads sleep instead of spinning
            // while the calling thread launches the next batch
            HIP_TRY(hipEventCreateWithFlags(&E.slot_ev[s],
                        hipEventDisableTiming | hipEventBlockingSync));
    }
    if(E.cs == nullptr)
        HIP_TRY(hipStreamCreateWithFlags(&E.cs, hipStreamNonBlocking));
    for(int s = 0; s < NSLOT; s++)
        E.slot_free[s] = true;
    E.used = 0;
    E.seg_used = 0;
    E.jobs.clear();
    E.next_job = 0;
    E.tasks.clear();
    E.ending = false;
    E.dispatch_done = false;
    E.error = STARSH_SUCCESS;
    E.any = false;
    trace_on = getenv("STARSH_HIP_COPY_TRACE") != nullptr;
    E.t_first = std::chrono::steady_clock::now();
    // workers: the cores left after the calling thread and the dispatcher
    unsigned nth = std::thread::hardware_concurrency();
    nth = nth < 4 ? 2 : (nth > 16 ? 14 : nth-2);
    E.dispatcher = std::thread(dispatcher_loop);
    for(unsigned i = 0; i < nth; i++)
        E.workers.emplace_back(worker_loop);
    E.running = true;
    return STARSH_SUCCESS;
}

static int queue_job(std::unique_ptr<Job> job);

int rsdd_out_append(const RsddOutput *out, int ntiles, const int *m,
        const int *n, const int *take, double *const *dst_u,
        double *const *dst_v)
{
    if(!E.running)
        return STARSH_UNKNOWN_ERROR;
    hipStream_t s = starsh_hip_stream();
    auto job = std::make_unique<Job>();
    job->lo = E.used;
    int64_t *seg = E.hseg+E.seg_used;
    int nseg = 0;
    int64_t at = E.used;
    for(int t = 0; t < ntiles; t++)
    {
        if(!take[t] || out->rank[t] <= 0)
            continue;
        const int64_t lu = (int64_t)m[t]*out->rank[t];
        const int64_t lv = (int64_t)n[t]*out->rank[t];
        if(at+lu+lv > E.arena_cap || E.seg_used+3*(nseg+2) > E.seg_cap)
        {
            fprintf(stderr, "STARSH HIP ERROR: factor arena too small\n");
            return STARSH_UNKNOWN_ERROR;
        }
        seg[3*nseg] = out->uoff[t];
        seg[3*nseg+1] = at;
        seg[3*nseg+2] = lu;
        job->segs.push_back(Seg{at, lu, dst_u[t]});
        nseg++;
        at += lu;
        seg[3*nseg] = out->voff[t];
        seg[3*nseg+1] = at;
        seg[3*nseg+2] = lv;
        job->segs.push_back(Seg{at, lv, dst_v[t]});
        nseg++;
        at += lv;
    }
    if(nseg == 0)
        return STARSH_SUCCESS;
    int64_t *dseg = E.dseg+E.seg_used;
    HIP_TRY(hipMemcpyAsync(dseg, seg, sizeof(int64_t)*3*(size_t)nseg,
                hipMemcpyHostToDevice, s));
    for(int s0 = 0; s0 < nseg; s0 += 65535)
    {
        const int cnt = std::min(nseg-s0, 65535);
        hipLaunchKernelGGL(arena_pack_kernel, dim3(16, cnt), dim3(256), 0, s,
                out->d_uv, E.arena, dseg+3*(int64_t)s0, cnt);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventCreateWithFlags(&job->ready, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(job->ready, s));
    E.seg_used += 3*(int64_t)nseg;
    E.used = at;
    job->hi = at;
    job->src = E.arena+job->lo;
    // arena offsets are relative to the job's range from here on
    for(auto &sg : job->segs)
        sg.aoff -= job->lo;
    return queue_job(std::move(job));
}

// Queues a job whose ready event has been recorded: its destination pages
// are touched while its data is sent.
static int queue_job(std::unique_ptr<Job> job)
{
    {
        std::lock_guard<std::mutex> g(E.mu);
        if(!E.any)
        {
            E.any = true;
            E.t_job0 = E.t_last = std::chrono::steady_clock::now();
        }
        // the destination pages are touched while the batch's factors are
        // packed and sent
        Job *jp = job.get();
        const int64_t len = jp->hi-jp->lo;
        trace_add('j', -1, jp->lo, len);
        jp->touched.assign((size_t)((len+TOUCH-1)/TOUCH), 0);
        for(int64_t a = 0; a < len; a += TOUCH)
            E.tasks.push_back(Task{-1, a, std::min(TOUCH, len-a), 0, jp});
        E.jobs.push_back(std::move(job));
    }
    E.cv_piece.notify_all();
    E.cv_job.notify_one();
    return STARSH_SUCCESS;
}

int rsdd_out_submit(const double *d_src, int64_t len, int nseg,
        const int64_t *off, const int64_t *seglen, double *const *dst)
{
    if(!E.running)
        return STARSH_UNKNOWN_ERROR;
    if(len <= 0 || nseg <= 0)
        return STARSH_SUCCESS;
    auto job = std::make_unique<Job>();
    job->src = d_src;
    job->lo = 0;
    job->hi = len;
    for(int q = 0; q < nseg; q++)
    {
        if(off[q] < 0 || seglen[q] < 0 || off[q]+seglen[q] > len ||
                (q > 0 && off[q] < off[q-1]+seglen[q-1]))
            return STARSH_WRONG_PARAMETER;
        job->segs.push_back(Seg{off[q], seglen[q], dst[q]});
    }
    HIP_TRY(hipEventCreateWithFlags(&job->ready, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(job->ready, starsh_hip_stream()));
    return queue_job(std::move(job));
}

static void trace_dump(void)
{
    const char *path = getenv("STARSH_HIP_COPY_TRACE");
    if(path == nullptr)
        return;
    FILE *f = fopen(path, "w");
    if(f == nullptr)
        return;
    for(const auto &r : trace)
        fprintf(f, "%c %d %lld %lld %.1f\n", r.kind, r.slot,
                (long long)r.off, (long long)r.len, r.us);
    fclose(f);
    trace.clear();
}

int rsdd_out_end(double *window_seconds, double *tail_seconds)
{
    if(!E.running)
        return STARSH_SUCCESS;
    const auto t0 = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> g(E.mu);
        E.ending = true;
    }
    E.cv_job.notify_all();
    E.dispatcher.join();
    for(auto &w : E.workers)
        w.join();
    E.workers.clear();
    const auto t1 = std::chrono::steady_clock::now();
    trace_add('e', -1, 0, 0);
    if(trace_on)
        trace_dump();
    for(auto &j : E.jobs)
        if(j->ready)
            (void)hipEventDestroy(j->ready);
    E.jobs.clear();
    E.running = false;
    if(E.any)
    {
        if(window_seconds)
            *window_seconds += std::chrono::duration<double>(E.t_last-
                    E.t_job0).count();
        if(tail_seconds)
            *tail_seconds += std::chrono::duration<double>(t1-t0).count();
    }
    return E.error;
}

void rsdd_out_release(void)
{
    if(E.running)
        return;
    if(E.arena)
        (void)hipFree(E.arena);
    E.arena = nullptr;
    E.arena_cap = 0;
    if(E.dseg)
        (void)hipFree(E.dseg);
    if(E.hseg)
        (void)hipHostFree(E.hseg);
    E.dseg = E.hseg = nullptr;
    E.seg_cap = 0;
}

extern "C" int starsh_hip_copy_option(int which, int value)
{
    if(E.running)
        return STARSH_UNKNOWN_ERROR;
    if(which == 0)
    {
        // slot size in MiB: 1 .. 64, a power of two
        if(value < 1 || value > 64 || (value & (value-1)) != 0)
            return STARSH_WRONG_PARAMETER;
        SLOT = (int64_t)value << 17;
        NSLOT = (int)(RING/SLOT);
        return STARSH_SUCCESS;
    }
    if(which == 1)
    {
        if(value != 0 && value != 1)
            return STARSH_WRONG_PARAMETER;
        use_lib_stream = value;
        return STARSH_SUCCESS;
    }
    return STARSH_WRONG_PARAMETER;
}
