// Orchestration of the plain FP64 randomized-SVD pipeline over a batch of
// tiles (src/backends/sequential/dense/drsdd.c:16-93 applied to every tile
// of src/backends/openmp/blrm/drsdd.c:94-145):
//   Y = A Omega; Q = orth(Y); Bt = A^T Q; truncate; Bt[:, :p] = Q2 R2;
//   Jacobi SVD of R2^T; rank; U = Q X_r; V = Bt X_r.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <thread>
#include <chrono>
#include <vector>
#include <utility>
#include "pipeline.h"
#include "rsdd.h"

namespace {

// Deflation thresholds of the Gram-Schmidt passes (relative to the column
// norm before projection, and to 1 for the second pass).
constexpr double TAU_FIRST = 1e-13;
// Largest core SVD (pipeline_core).
constexpr int CORE_PMAX = 1024;
constexpr double TAU_SECOND = 1e-8;
// Truncation of B's rows: discarded tail at most (ETA*tol)^2 |B|_F^2
// (RsddInput.eta overrides it).
constexpr double ETA = 1e-5;

#define TRY(expr) do { int i_ = (expr); if(i_ != STARSH_SUCCESS) return i_; } \
    while(0)

// Grow-only device (or pinned host) buffer kept between calls, so that
// repeated compressions do not pay for allocating tens of GB each time.
struct Pool
{
    void *p = nullptr;
    size_t bytes = 0;
    bool host = false;
    // allocations are rounded up to this to absorb small variations
    // between calls
    size_t unit = (size_t)64 << 20;
    int reserve(size_t want)
    {
        if(want <= bytes && p)
            return STARSH_SUCCESS;
        release();
        size_t b = (want+unit-1)/unit*unit;
        hipError_t e = host ? hipHostMalloc(&p, b, hipHostMallocDefault) :
            hipMalloc(&p, b);
        if(e != hipSuccess)
        {
            p = nullptr;
            fprintf(stderr, "STARSH HIP ERROR: allocation of %zu bytes "
                    "failed\n", b);
            return STARSH_MALLOC_ERROR;
        }
        bytes = b;
        return STARSH_SUCCESS;
    }
    void release()
    {
        if(p)
        {
            (void)hipStreamSynchronize(starsh_hip_stream());
            if(host)
                (void)hipHostFree(p);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
};

Pool ws_pool, uv_pool, pack_pool, stage_pool{nullptr, 0, true};
// The small per-batch buffers (tile descriptors, kept-column flags, norms,
// int8 Omega and digit planes, counters, R2 partials, selected columns,
// pivots), kept too: a hipMalloc / hipFree pair per buffer and batch cost
// host time on every call, and hipFree waits for the device.
constexpr size_t SMALL_UNIT = (size_t)1 << 20;
Pool tiles_pool{nullptr, 0, false, SMALL_UNIT},
     keep_pool{nullptr, 0, false, SMALL_UNIT},
     norms_pool{nullptr, 0, false, SMALL_UNIT},
     om8_pool{nullptr, 0, false, SMALL_UNIT},
     active_pool{nullptr, 0, false, SMALL_UNIT},
     qd_pool{nullptr, 0, false, SMALL_UNIT},
     part_pool{nullptr, 0, false, SMALL_UNIT},
     sel_pool{nullptr, 0, false, SMALL_UNIT},
     piv_pool{nullptr, 0, false, SMALL_UNIT};

// Host view of the device counter of tiles still in a BCGS2 run.  The
// counter is copied to pinned memory after every panel; the host reads it
// LAG panels later (so the stream always holds queued work) and stops
// launching once no tile is left.
constexpr int LAG = 3;
constexpr int NSLOT = 4;

struct ActiveProbe
{
    int *h = nullptr;          // pinned host slots
    hipEvent_t ev[NSLOT] = {};
    int init()
    {
        if(h)
            return STARSH_SUCCESS;
        HIP_TRY(hipHostMalloc((void **)&h, sizeof(int)*NSLOT,
                    hipHostMallocDefault));
        for(int e = 0; e < NSLOT; e++)
            HIP_TRY(hipEventCreateWithFlags(&ev[e], hipEventDisableTiming));
        return STARSH_SUCCESS;
    }
};

ActiveProbe probe;

// Panel width of the block Gram-Schmidt: 0 (by batch, default), 4, or 8
// (tiles of at most 2048 rows); set by starsh_hip_set_option("qr_panel").
// A 4-column panel keeps two tiles resident per CU (512 on the chip), an
// 8-column one a single tile, with half the launches and half the re-reads
// of earlier columns.  Batches of up to PANEL8_TILES tiles are done in at
// most three rounds of 4-column workgroups and take them (C2, 820 tiles:
// 10.1 ms against 11.7 ms with 8 columns); larger ones take 8 columns (C3,
// 2080 tiles: 1.61 s against 1.72 s in the plain form).
int qr_panel_width = 0;
// The width chosen by the autotuner for the current call (0 = by batch);
// the option above wins.
int qr_panel_tuned = 0;
constexpr int PANEL8_TILES = 1536;

// Projection of the mixed form on int8 digit planes (1, default) or FP64
// MFMA (0); set by starsh_hip_set_option("proj_int8").
int proj_int8 = 1;

// Gram-Schmidt panels as one captured hipGraph (option "qr_graph" 1)
// instead of host launches pruned by the lagged probe: every panel up to
// the batch's width is in the graph (finished tiles return at once), and
// the executable graph is kept and replayed while its arguments (the
// pooled buffers, sizes and thresholds) are unchanged, so a repeated
// compression of one problem captures once.
int qr_graph = 0;

struct PanelGraph
{
    struct Key
    {
        const void *tiles, *ws, *keep, *active;
        int ntiles, which, kmax, width_max, rows_max, pb;
        double tau_first;
        bool operator==(const Key &o) const
        {
            return tiles == o.tiles && ws == o.ws && keep == o.keep &&
                active == o.active && ntiles == o.ntiles &&
                which == o.which && kmax == o.kmax &&
                width_max == o.width_max && rows_max == o.rows_max &&
                pb == o.pb && tau_first == o.tau_first;
        }
    } key{};
    hipGraphExec_t exec = nullptr;
    void reset()
    {
        if(exec)
            (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
};
PanelGraph panel_graphs[2];      // which = 0, 1

int bcgs2(TileDesc *d_tiles, int ntiles, int which, double *ws, int kmax,
        int width_max, int rows_max, int *d_keep, double tau_first,
        int *d_active, hipStream_t s)
{
    const int pw = qr_panel_width ? qr_panel_width : qr_panel_tuned;
    const bool wide = pw == 8 || (pw == 0 && ntiles > PANEL8_TILES);
    const int pb = rows_max <= 2048 && wide ? 8 : 4;
    if(qr_graph)
    {
        PanelGraph &pg = panel_graphs[which ? 1 : 0];
        const PanelGraph::Key key{d_tiles, ws, d_keep, d_active, ntiles,
            which, kmax, width_max, rows_max, pb, tau_first};
        TRY(pipeline_count_active(d_tiles, ntiles, which, d_active, s));
        // captured and launched on a stream of its own, ordered after and
        // before the library stream's work by events: the copy-back threads
        // queue transfers on the library stream, and those run at the
        // library stream's rate only if no graph is launched on it
        static hipStream_t cs = nullptr;
        static hipEvent_t before = nullptr, after = nullptr;
        if(cs == nullptr)
        {
            HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&before, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&after, hipEventDisableTiming));
        }
        if(!(pg.exec && pg.key == key))
        {
            pg.reset();
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(cs,
                        hipStreamCaptureModeThreadLocal));
            int info = STARSH_SUCCESS;
            for(int j0 = 0; info == STARSH_SUCCESS && j0 < width_max;
                    j0 += pb)
                info = pipeline_panel(d_tiles, ntiles, which, ws, j0, pb,
                        kmax, tau_first, TAU_SECOND, d_keep, rows_max,
                        d_active, cs);
            hipError_t e = hipStreamEndCapture(cs, &graph);
            if(info != STARSH_SUCCESS || e != hipSuccess)
            {
                if(graph)
                    (void)hipGraphDestroy(graph);
                return info != STARSH_SUCCESS ? info : STARSH_UNKNOWN_ERROR;
            }
            e = hipGraphInstantiate(&pg.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess)
            {
                pg.exec = nullptr;
                return STARSH_UNKNOWN_ERROR;
            }
            pg.key = key;
        }
        HIP_TRY(hipEventRecord(before, s));
        HIP_TRY(hipStreamWaitEvent(cs, before, 0));
        HIP_TRY(hipGraphLaunch(pg.exec, cs));
        HIP_TRY(hipEventRecord(after, cs));
        HIP_TRY(hipStreamWaitEvent(s, after, 0));
        return STARSH_SUCCESS;
    }
    TRY(probe.init());
    TRY(pipeline_count_active(d_tiles, ntiles, which, d_active, s));
    int panel = 0;
    for(int j0 = 0; j0 < width_max; j0 += pb, panel++)
    {
        if(panel >= LAG)
        {
            const int slot = (panel-LAG) % NSLOT;
            HIP_TRY(hipEventSynchronize(probe.ev[slot]));
            if(probe.h[slot] == 0)
                break;
        }
        TRY(pipeline_panel(d_tiles, ntiles, which, ws, j0, pb, kmax,
                    tau_first, TAU_SECOND, d_keep, rows_max, d_active, s));
        const int slot = panel % NSLOT;
        HIP_TRY(hipMemcpyAsync(probe.h+slot, d_active, sizeof(int),
                    hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(probe.ev[slot], s));
    }
    return STARSH_SUCCESS;
}

// Diagnostic: when STARSH_HIP_DEBUG_DUMP names a directory, the buffers of
// tile 0 are written there after each stage (raw doubles, column-major).
void dump(const char *name, const double *d_src, int64_t count, hipStream_t s)
{
    const char *dir = getenv("STARSH_HIP_DEBUG_DUMP");
    if(dir == nullptr || count <= 0)
        return;
    std::vector<double> h(count);
    if(hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpy(h.data(), d_src, sizeof(double)*count,
                hipMemcpyDeviceToHost) != hipSuccess)
        return;
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", dir, name);
    FILE *f = fopen(path, "wb");
    if(f)
    {
        fwrite(h.data(), sizeof(double), count, f);
        fclose(f);
    }
}

} // namespace

namespace {

// Gathers segments of a device buffer into one packed device buffer.
__global__ __launch_bounds__(256) void pack_kernel(const double *src,
        double *dst, const int64_t *seg, int nseg)
{
    // seg[3 s .. 3 s + 2] = source offset, packed offset, length
    const int s = blockIdx.y;
    if(s >= nseg)
        return;
    const int64_t so = seg[3*s], po = seg[3*s+1], len = seg[3*s+2];
    for(int64_t e = (int64_t)blockIdx.x*blockDim.x+threadIdx.x; e < len;
            e += (int64_t)gridDim.x*blockDim.x)
        dst[po+e] = src[so+e];
}


// Copies the singular values (k per tile, from the workspace) to
// S_out + soff[t], through one packed device-to-host transfer.
int copy_segments(const double *d_ws, const std::vector<TileDesc> &h,
        double *S_out, const int64_t *soff, hipStream_t s)
{
    const int nt = (int)h.size();
    std::vector<int64_t> seg(3*(size_t)nt);
    int64_t total = 0;
    for(int t = 0; t < nt; t++)
    {
        seg[3*t] = h[t].s_off;
        seg[3*t+1] = total;
        seg[3*t+2] = h[t].k;
        total += h[t].k;
    }
    TRY(pack_pool.reserve(sizeof(double)*total+sizeof(int64_t)*seg.size()));
    TRY(stage_pool.reserve(sizeof(double)*total));
    double *d_pack = (double *)pack_pool.p;
    int64_t *d_seg = (int64_t *)(d_pack+total);
    HIP_TRY(hipMemcpyAsync(d_seg, seg.data(), sizeof(int64_t)*seg.size(),
                hipMemcpyHostToDevice, s));
    for(int s0 = 0; s0 < nt; s0 += 65535)
    {
        const int cnt = std::min(nt-s0, 65535);
        hipLaunchKernelGGL(pack_kernel, dim3(1, cnt), dim3(256), 0, s, d_ws,
                d_pack, d_seg+3*(int64_t)s0, cnt);
        HIP_TRY(hipGetLastError());
    }
    double *hs = (double *)stage_pool.p;
    HIP_TRY(hipMemcpyAsync(hs, d_pack, sizeof(double)*total,
                hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for(int t = 0; t < nt; t++)
        memcpy(S_out+soff[t], hs+seg[3*t+1], sizeof(double)*h[t].k);
    return STARSH_SUCCESS;
}

} // namespace

// Runs the pipeline on ntiles tiles.  Either the kernel (kd, points) or
// explicit tiles (d_a, a_off) define the tile contents.  Results: ranks in
// rank_out, U/V in device memory (*d_uv, offsets in uoff/voff, allocated
// here and owned by the caller), singular values (k per tile) copied to
// S_out + soff[t] when S_out is not NULL.
int rsdd_batch(const RsddInput *in, RsddOutput *out)
{
    hipStream_t s = starsh_hip_stream();
    const int nt = in->ntiles;
    out->d_uv = nullptr;
    if(nt <= 0)
        return STARSH_SUCCESS;
    // problems of the batch
    RsddProblem single;
    single.kd = in->kd;
    single.d_xr = in->d_xr;
    single.d_xc = in->d_xc;
    single.nr = in->nr;
    single.nc = in->nc;
    single.maxrank = in->maxrank;
    single.oversample = in->oversample;
    single.tol = in->tol;
    const RsddProblem *probs = in->probs ? in->probs : &single;
    const int nprob = in->probs ? in->nprob : 1;
    std::vector<KernelParams> kps(nprob);
    for(int q = 0; q < nprob; q++)
    {
        memset(&kps[q], 0, sizeof(KernelParams));
        if(probs[q].kd)
            kps[q] = make_params(probs[q].kd);
        else
            kps[q].ndim = 1;
    }
    // the tiles of one batch share a kernel family and dimension
    const int q0 = in->probs && in->prob ? in->prob[0] : 0;
    if(q0 < 0 || q0 >= nprob)
        return STARSH_WRONG_PARAMETER;
    std::vector<TileDesc> h(nt);
    std::map<std::pair<int, int>, int64_t> omegas;
    int64_t off = 0;
    int kmax = 0, mmax = 0, nmax = 0;
    for(int t = 0; t < nt; t++)
    {
        TileDesc &d = h[t];
        memset(&d, 0, sizeof(d));
        const int q = in->probs && in->prob ? in->prob[t] : 0;
        if(q < 0 || q >= nprob || kps[q].ndim != kps[q0].ndim ||
                kps[q].family != kps[q0].family)
            return STARSH_WRONG_PARAMETER;
        const RsddProblem &pr = probs[q];
        d.xr = pr.d_xr;
        d.xc = pr.d_xc;
        d.nr = pr.nr;
        d.nc = pr.nc;
        d.kp = kps[q];
        d.tol = pr.tol;
        d.maxrank = pr.maxrank;
        d.m = in->m[t];
        d.n = in->n[t];
        int mn = d.m < d.n ? d.m : d.n;
        d.k = in->k ? in->k[t] : pr.maxrank+pr.oversample;
        if(d.k > mn)
            d.k = mn;
        if(d.k < 1)
            d.k = 1;
        d.gen = in->d_a == nullptr;
        d.rule = in->rule;
        d.sel_off = in->sel_off ? in->sel_off[t] : 0;
        d.piv_k = in->piv_k ? in->piv_k[t] : 0;
        d.piv_off = d.piv_k > 0 && in->piv_off ? in->piv_off[t] : -1;
        d.r0 = in->r0 ? in->r0[t] : 0;
        d.c0 = in->c0 ? in->c0[t] : 0;
        d.a_off = in->a_off ? in->a_off[t] : 0;
        kmax = d.k > kmax ? d.k : kmax;
        mmax = d.m > mmax ? d.m : mmax;
        nmax = d.n > nmax ? d.n : nmax;
    }
    // Workspace layout: all Omega first, then per-tile regions.  The int8
    // sketch keeps its integer Omega in a separate byte buffer.
    int64_t off8 = 0;
    const bool select = (in->rule & RSDD_RULE_SELECT) != 0;
    if(select && (in->sel == nullptr || in->sel_off == nullptr ||
                in->sketch_int8))
        return STARSH_WRONG_PARAMETER;
    for(int t = 0; t < nt && !select; t++)
    {
        auto key = std::make_pair(h[t].n, h[t].k);
        if(omegas.find(key) == omegas.end())
        {
            if(in->sketch_int8)
            {
                omegas[key] = off8;
                off8 += (int64_t)((h[t].n+63)/64*64)*h[t].k;
            }
            else
            {
                omegas[key] = off;
                off += (int64_t)h[t].n*h[t].k;
            }
        }
        if(in->sketch_int8)
            h[t].om8_off = omegas[key];
        else
            h[t].om_off = omegas[key];
    }
    // int8 digit planes of Q for the projection of the mixed form
    const bool proj8 = in->sketch_int8 && proj_int8;
    int64_t offq = 0;
    if(proj8)
        for(int t = 0; t < nt; t++)
        {
            h[t].qd_off = offq;
            offq += pipeline_qdigits_bytes(h[t].m, h[t].k);
        }
    int64_t uvoff = 0;
    for(int t = 0; t < nt; t++)
    {
        TileDesc &d = h[t];
        int64_t m = d.m, n = d.n, k = d.k;
        int64_t rmax = d.maxrank < k ? d.maxrank : k;
        d.q_off = off; off += m*k;
        d.bt_off = off; off += n*k;
        d.q2_off = off; off += n*k;
        d.r_off = off; off += 2*k*k;
        d.w_off = off; off += 2*k*k;
        d.s_off = off; off += k;
        d.u_off = uvoff; uvoff += m*rmax;
        d.v_off = uvoff; uvoff += n*rmax;
        out->uoff[t] = d.u_off;
        out->voff[t] = d.v_off;
    }
    if(in->rows_out)
        for(int t = 0; t < nt; t++)
            in->rows_out[t] = h[t].k;
    // selected columns (selection sketch) and column pivots (RRQR)
    int64_t nsel = 0, npiv = 0;
    for(int t = 0; t < nt; t++)
    {
        if(select)
            nsel = std::max<int64_t>(nsel, h[t].sel_off+h[t].k);
        if(h[t].piv_off >= 0)
            npiv = std::max<int64_t>(npiv, h[t].piv_off+h[t].piv_k);
    }
    // part: split-K partials of R2, kept until the batch has synchronised
    Pool &tiles = tiles_pool, &keep = keep_pool, &norms = norms_pool,
         &om8 = om8_pool, &active = active_pool, &qd = qd_pool,
         &part = part_pool, &selbuf = sel_pool, &pivbuf = piv_pool;
    if(nsel > 0)
    {
        TRY(selbuf.reserve(sizeof(int)*(size_t)nsel));
        HIP_TRY(hipMemcpyAsync(selbuf.p, in->sel, sizeof(int)*(size_t)nsel,
                    hipMemcpyHostToDevice, s));
    }
    if(npiv > 0)
    {
        if(in->piv == nullptr)
            return STARSH_WRONG_PARAMETER;
        TRY(pivbuf.reserve(sizeof(int)*(size_t)npiv));
    }
    TRY(ws_pool.reserve(sizeof(double)*off));
    if(in->sketch_int8)
        TRY(om8.reserve(off8));
    if(proj8)
        TRY(qd.reserve(offq));
    TRY(tiles.reserve(sizeof(TileDesc)*nt));
    TRY(keep.reserve(sizeof(int)*(size_t)nt*kmax));
    TRY(norms.reserve(sizeof(double)*(size_t)nt*kmax));
    TRY(uv_pool.reserve(sizeof(double)*(uvoff > 0 ? uvoff : 1)));
    TRY(active.reserve(sizeof(int)));
    TileDesc *d_tiles = (TileDesc *)tiles.p;
    double *d_ws = (double *)ws_pool.p;
    HIP_TRY(hipMemcpyAsync(d_tiles, h.data(), sizeof(TileDesc)*nt,
                hipMemcpyHostToDevice, s));
    for(auto &o : omegas)
    {
        if(in->sketch_int8)
            TRY(pipeline_omega_i8((signed char *)om8.p+o.second,
                        o.first.first, o.first.second, s));
        else
            TRY(starsh_hip_omega(d_ws+o.second, o.first.first,
                        o.first.second));
    }

    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.ws = d_ws;
    g.out = (double *)uv_pool.p;
    g.abase = in->d_a;
    g.ndim = kps[q0].ndim;
    g.family = kps[q0].family;
    for(const auto &kp : kps)
        g.noisy |= kp.noise != 0.0;

    // destroyed on every return path
    struct Events
    {
        hipEvent_t e[6] = {nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr};
        ~Events()
        {
            for(auto x : e)
                if(x)
                    (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t *ev = evs.e;
    for(int e = 0; e < 6; e++)
        HIP_TRY(hipEventCreate(&ev[e]));
    HIP_TRY(hipEventRecord(ev[0], s));
    // 1. sketch
    HIP_TRY(hipEventRecord(ev[2], s));
    if(select)
        TRY(pipeline_select(d_tiles, nt, mmax, kmax, g.ndim, d_ws,
                    (const int *)selbuf.p, in->d_a, s));
    else if(in->sketch_int8)
    {
        if(in->d_a != nullptr || probs[q0].kd == nullptr)
            return STARSH_WRONG_PARAMETER;
        TRY(pipeline_sketch_i8(d_tiles, nt, mmax, kmax, g,
                    (const signed char *)om8.p, s));
    }
    else
    {
        if(in->d_a)
            TRY(pipeline_gemm(GM_SKETCH_MEM, d_tiles, nt, mmax, kmax, g, s));
        else
            TRY(pipeline_gen_gemm(0, d_tiles, nt, mmax, kmax, g, s));
    }
    HIP_TRY(hipEventRecord(ev[3], s));
    dump("omega", d_ws+h[0].om_off, (int64_t)h[0].n*h[0].k, s);
    dump("Y", d_ws+h[0].q_off, (int64_t)h[0].m*h[0].k, s);
    // 2. orthonormal basis with deflation, kept columns to the front
    // columns of panels that are never launched count as dropped
    HIP_TRY(hipMemsetAsync(keep.p, 0, sizeof(int)*(size_t)nt*kmax, s));
    TRY(bcgs2(d_tiles, nt, 0, d_ws, kmax, kmax, mmax, (int *)keep.p,
                in->tau > 0 ? in->tau : TAU_FIRST, (int *)active.p, s));
    dump("Qfull", d_ws+h[0].q_off, (int64_t)h[0].m*h[0].k, s);
    TRY(pipeline_compact(d_tiles, nt, d_ws, (int *)keep.p, kmax, s));
    dump("Q", d_ws+h[0].q_off, (int64_t)h[0].m*h[0].k, s);
    // 3. projection Bt = A^T Q
    HIP_TRY(hipEventRecord(ev[4], s));
    if(in->d_a)
        TRY(pipeline_gemm(GM_PROJ_MEM, d_tiles, nt, nmax, kmax, g, s));
    else if(proj8)
    {
        TRY(pipeline_qdigits(d_tiles, nt, kmax, d_ws, (signed char *)qd.p,
                    s));
        // the projection kernel alone
        HIP_TRY(hipEventRecord(ev[4], s));
        TRY(pipeline_proj_i8(d_tiles, nt, nmax, kmax, g,
                    (const signed char *)qd.p, s));
    }
    else
        TRY(pipeline_gen_gemm(1, d_tiles, nt, nmax, kmax, g, s));
    HIP_TRY(hipEventRecord(ev[5], s));
    dump("Bt", d_ws+h[0].bt_off, (int64_t)h[0].n*h[0].k, s);
    // 4. truncation of B's rows
    TRY(pipeline_colnorm(d_tiles, nt, 2, d_ws, (double *)norms.p, kmax, s));
    TRY(pipeline_truncate(d_tiles, nt, d_ws, (double *)norms.p, kmax,
                in->eta > 0 ? in->eta : ETA, nmax, s));
    HIP_TRY(hipMemcpyAsync(h.data(), d_tiles, sizeof(TileDesc)*nt,
                hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // The core SVD takes p <= CORE_PMAX.  A tile whose truncated B is
    // larger (only in the full-width rounds of the exact engines) keeps its
    // first CORE_PMAX rows; the squared norms of the rows dropped are added
    // to its tail, as for the truncation itself, so the rank rule stays
    // conservative (flag 8).
    int pmax = 0;
    bool capped = false;
    for(int t = 0; t < nt; t++)
    {
        if(h[t].p > CORE_PMAX)
        {
            std::vector<double> nb(h[t].p-CORE_PMAX);
            HIP_TRY(hipMemcpy(nb.data(), (double *)norms.p+(int64_t)t*kmax+
                        CORE_PMAX, sizeof(double)*nb.size(),
                        hipMemcpyDeviceToHost));
            double extra = 0.0;
            for(size_t c = nb.size(); c > 0; c--)
                extra += nb[c-1];
            h[t].tail2 += extra;
            h[t].p = CORE_PMAX;
            h[t].flags |= 8;
            capped = true;
        }
        pmax = h[t].p > pmax ? h[t].p : pmax;
    }
    if(capped)
        HIP_TRY(hipMemcpyAsync(d_tiles, h.data(), sizeof(TileDesc)*nt,
                    hipMemcpyHostToDevice, s));
    // 5. Bt[:, :p] = Q2 R2
    if(pmax > 0)
    {
        TRY(bcgs2(d_tiles, nt, 1, d_ws, kmax, pmax, nmax, (int *)keep.p,
                    TAU_FIRST, (int *)active.p, s));
        dump("Q2", d_ws+h[0].q2_off, (int64_t)h[0].n*h[0].p, s);
        const int nz = pipeline_r2_split(nt, pmax, nmax);
        if(nz > 1)
            TRY(part.reserve(sizeof(double)*(size_t)nt*nz*pmax*pmax));
        TRY(pipeline_r2(d_tiles, nt, pmax, nz, (double *)part.p, g, s));
        dump("R2", d_ws+h[0].r_off, (int64_t)h[0].k*h[0].k, s);
    }
    // 6. core SVD and rank
    TRY(pipeline_core(d_tiles, nt, d_ws, pmax, s));
    HIP_TRY(hipMemcpyAsync(h.data(), d_tiles, sizeof(TileDesc)*nt,
                hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int rmax = 0;
    for(int t = 0; t < nt; t++)
        rmax = h[t].rank > rmax ? h[t].rank : rmax;
    // 7. factors
    if(rmax > 0)
    {
        TRY(pipeline_gemm(GM_FORM_U, d_tiles, nt, mmax, rmax, g, s));
        TRY(pipeline_gemm(GM_FORM_V, d_tiles, nt, nmax, rmax, g, s));
    }
    // 8. column pivots of B for the tiles that need them (Bt is no longer
    // used and is overwritten)
    if(npiv > 0)
    {
        TRY(pipeline_qrcp(d_tiles, nt, nmax, kmax, d_ws, (int *)pivbuf.p,
                    s));
        HIP_TRY(hipMemcpyAsync(in->piv, pivbuf.p, sizeof(int)*(size_t)npiv,
                    hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipEventRecord(ev[1], s));
    if(in->S_out)
        TRY(copy_segments(d_ws, h, in->S_out, in->soff, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0, ms_sk = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms_sk, ev[2], ev[3]));
    float ms_pr = 0;
    HIP_TRY(hipEventElapsedTime(&ms_pr, ev[4], ev[5]));
    out->proj_seconds = ms_pr*1e-3;
    out->device_seconds = ms*1e-3;
    out->sketch_seconds = ms_sk*1e-3;
    for(int t = 0; t < nt; t++)
    {
        out->rank[t] = h[t].rank;
        out->p[t] = h[t].p;
        out->kc[t] = h[t].kc;
        out->flags[t] = h[t].flags;
        if(in->piv_done)
            in->piv_done[t] = h[t].piv_off >= 0 && h[t].kc > h[t].piv_k;
    }
    out->d_uv = (double *)uv_pool.p;
    return STARSH_SUCCESS;
}


// Frees the buffers kept between calls.
void rsdd_release_pools(void)
{
    ws_pool.release();
    uv_pool.release();
    pack_pool.release();
    stage_pool.release();
    for(Pool *q : {&tiles_pool, &keep_pool, &norms_pool, &om8_pool,
                &active_pool, &qd_pool, &part_pool, &sel_pool, &piv_pool})
        q->release();
    for(PanelGraph &pg : panel_graphs)
        pg.reset();
    rsdd_out_release();
}

extern "C" int starsh_hip_proj_int8(int on)
{
    if(on != 0 && on != 1)
        return STARSH_WRONG_PARAMETER;
    proj_int8 = on;
    return STARSH_SUCCESS;
}

extern "C" int hip_set_tuned_panel(int width)
{
    if(width != 0 && width != 4 && width != 8)
        return STARSH_WRONG_PARAMETER;
    qr_panel_tuned = width;
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_qr_graph(int on)
{
    if(on != 0 && on != 1)
        return STARSH_WRONG_PARAMETER;
    qr_graph = on;
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_qr_panel(int width)
{
    if(width != 0 && width != 4 && width != 8)
        return STARSH_WRONG_PARAMETER;
    qr_panel_width = width;
    return STARSH_SUCCESS;
}
