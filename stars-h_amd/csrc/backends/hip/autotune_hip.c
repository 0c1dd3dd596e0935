/* Autotuner of the HIP compression forms and its schedule table (no
 * reference counterpart: the reference's OpenMP/MPI/StarPU drivers run one
 * fixed schedule, src/backends/openmp/blrm/drsdd.c:94-145).
 *
 * What is tuned, per problem (one compression per trial, device time from
 * HIP events, every trial on the format as passed):
 *   plain form  the column block and row groups of the generated-tile
 *               sketch GEMM and of the projection GEMM, then the panel
 *               width of the block Gram-Schmidt;
 *   mixed form  the first rank bucket (sketch width 32..512), then the
 *               bucket ladder (with or without a 384-wide bucket), then a
 *               per-tile schedule (the bucket each far-field tile finished
 *               in, or "false far-field"), kept only if it is faster, then
 *               the Gram-Schmidt panel width.
 * A batch (starsh_blrm_approximate_batch) is tuned as one: its problems
 * share one set of knobs.
 *
 * The table, per problem signature (form, kernel and its parameters, point
 * count, largest cluster, maxrank, oversample, tolerance and a hash of the
 * point coordinates), holds the chosen knobs and the per-tile schedule.  A
 * compression looks its problems up: knobs set by starsh_hip_set_option
 * win over the table's, which win over the built-in choices by size.
 * Entries also go to the text file named by STARSH_HIP_SCHEDULE (read on
 * first use, appended to by the tuner), one line per entry:
 *   family ndim beta nu noise sigma wave_k diag n nb maxrank os tol form
 *   first ladder panel sketch_bn sketch_rw proj_bn proj_rw geom ntile tiles
 * tiles: one character per far-field tile, '0'-'4' ('0'-'5' on ladder 1) a
 * bucket of the entry's ladder, 'd' false far-field; "-" when there is no
 * per-tile schedule. */
#include <math.h>
#include "common.h"
#include "kernels_host.h"
#include "rsdd.h"
#include "hip_driver.h"

/* Buckets of ladder 0, on which the first bucket is tuned (ladder 1 has
 * one more, 384 wide, below the last). */
#define NBUCKETS 5
/* First bucket without a table entry: width 64.  On every workload
 * measured (2D sqexp and exp, 3D Matern, Coulomb and sin(kr)/r) width 64
 * beat width 32 as the first bucket. */
#define DEFAULT_FIRST 1

/* First bucket forced by starsh_hip_set_option("mixed_first_bucket"), or
 * -1 for the table. */
static int forced_bucket = -1;
/* Per-tile schedules used by compressions (option "tile_schedule"). */
static int use_tile_schedule = 1;

int hip_mixed_set_first_bucket(int b)
{
    if(b < -1 || b >= NBUCKETS)
        return STARSH_WRONG_PARAMETER;
    forced_bucket = b;
    return STARSH_SUCCESS;
}

int hip_mixed_set_tile_schedule(int on)
{
    if(on != 0 && on != 1)
        return STARSH_WRONG_PARAMETER;
    use_tile_schedule = on;
    return STARSH_SUCCESS;
}

typedef struct sched_entry
{
    starsh_hip_kernel_desc kd;
    long long n;
    int nb, maxrank, os;
    double tol;
    int form;               /* 0 plain, 1 mixed */
    int first;              /* mixed: first bucket (of `ladder`) */
    int ladder;             /* mixed: bucket ladder (0, 1) */
    int panel;              /* Gram-Schmidt panel width, 0 = by batch */
    int gemm_bn[2], gemm_rw[2];     /* [0] sketch, [1] projection */
    unsigned long long geom;
    int ntile;
    signed char *tiles;     /* malloc'd, or NULL */
} sched_entry;

/* FNV-1a hash of the coordinates of a problem's row points. */
static unsigned long long points_hash(const STARSH_blrf *F)
{
    starsh_hip_kernel_desc kd;
    const double *x = NULL;
    STARSH_int cnt = 0;
    unsigned long long h = 1469598103934665603ull;
    if(starsh_kernel_describe(F->problem->kernel, F->problem->row_data, &kd,
                &x, &cnt) != 0 || x == NULL)
        return h;
    const unsigned char *b = (const unsigned char *)x;
    const size_t nbytes = sizeof(double)*(size_t)cnt*(size_t)kd.ndim;
    for(size_t i = 0; i < nbytes; i++)
    {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

#define SCHED_MAX 256
static sched_entry sched[SCHED_MAX];
static int nsched = 0, sched_loaded = 0;

static void sched_key(int form, const hip_problem *hp, const STARSH_blrf *F,
        int maxrank, double tol, sched_entry *e)
{
    memset(e, 0, sizeof(*e));
    e->kd = hp->kd;
    e->n = hp->nr;
    int nb = 0;
    for(STARSH_int i = 0; i < F->row_cluster->nblocks; i++)
        if(F->row_cluster->size[i] > nb)
            nb = (int)F->row_cluster->size[i];
    e->nb = nb;
    e->maxrank = maxrank;
    e->os = hp->oversample;
    e->tol = tol;
    e->form = form;
    e->first = -1;
    e->geom = points_hash(F);
}

/* Same geometry: everything but the kernel's parameters. */
static int sched_same_geom(const sched_entry *a, const sched_entry *b)
{
    return a->form == b->form && a->kd.family == b->kd.family &&
        a->kd.ndim == b->kd.ndim && a->n == b->n && a->nb == b->nb &&
        a->maxrank == b->maxrank && a->os == b->os && a->tol == b->tol &&
        a->geom == b->geom;
}

static int sched_same(const sched_entry *a, const sched_entry *b)
{
    const starsh_hip_kernel_desc *x = &a->kd, *y = &b->kd;
    return sched_same_geom(a, b) && x->beta == y->beta && x->nu == y->nu &&
        x->noise == y->noise && x->sigma == y->sigma &&
        x->wave_k == y->wave_k && x->diag == y->diag;
}

/* Stores e, taking over its tile array. */
static void sched_put(sched_entry *e)
{
    for(int i = 0; i < nsched; i++)
        if(sched_same(&sched[i], e))
        {
            free(sched[i].tiles);
            sched[i] = *e;
            e->tiles = NULL;
            return;
        }
    if(nsched < SCHED_MAX)
        sched[nsched++] = *e;
    else
        free(e->tiles);
    e->tiles = NULL;
}

static int valid_knobs(const sched_entry *e)
{
    int ok = (e->form == 0 || e->form == 1) &&
        (e->ladder == 0 || e->ladder == 1) && e->first >= -1 &&
        e->first < hip_mixed_nbuckets(e->ladder) && (e->panel == 0 ||
                e->panel == 4 || e->panel == 8);
    for(int j = 0; j < 2; j++)
        ok = ok && (e->gemm_bn[j] == 0 || e->gemm_bn[j] == 64 ||
                e->gemm_bn[j] == 128 || e->gemm_bn[j] == 256) &&
            (e->gemm_rw[j] == 0 || e->gemm_rw[j] == 1 || e->gemm_rw[j] == 2);
    return ok;
}

static void sched_load(void)
{
    if(sched_loaded)
        return;
    sched_loaded = 1;
    const char *path = getenv("STARSH_HIP_SCHEDULE");
    FILE *f = path ? fopen(path, "r") : NULL;
    if(f == NULL)
        return;
    sched_entry e;
    memset(&e, 0, sizeof(e));
    while(fscanf(f, "%d %d %lg %lg %lg %lg %lg %lg %lld %d %d %d %lg %d %d "
                "%d %d %d %d %d %d %llu %d", &e.kd.family, &e.kd.ndim,
                &e.kd.beta, &e.kd.nu, &e.kd.noise, &e.kd.sigma, &e.kd.wave_k,
                &e.kd.diag, &e.n, &e.nb, &e.maxrank, &e.os, &e.tol, &e.form,
                &e.first, &e.ladder, &e.panel, &e.gemm_bn[0], &e.gemm_rw[0],
                &e.gemm_bn[1], &e.gemm_rw[1], &e.geom, &e.ntile) == 23)
    {
        e.tiles = NULL;
        int ok = valid_knobs(&e) && e.ntile >= 0 && e.ntile < (1 << 28);
        if(ok && e.ntile > 0)
        {
            char *str = malloc((size_t)e.ntile+1);
            e.tiles = malloc((size_t)e.ntile);
            char fmt[32];
            snprintf(fmt, sizeof(fmt), " %%%ds", e.ntile);
            ok = str && e.tiles && fscanf(f, fmt, str) == 1 &&
                (int)strlen(str) == e.ntile;
            for(int i = 0; ok && i < e.ntile; i++)
            {
                if(str[i] == 'd')
                    e.tiles[i] = -2;
                else if(str[i] >= '0' &&
                        str[i] < '0'+hip_mixed_nbuckets(e.ladder))
                    e.tiles[i] = (signed char)(str[i]-'0');
                else
                    ok = 0;
            }
            free(str);
        }
        else
        {
            char dash[8];
            ok = ok && fscanf(f, " %7s", dash) == 1;
            e.ntile = 0;
        }
        if(!ok)
        {
            free(e.tiles);
            break;
        }
        sched_put(&e);
    }
    fclose(f);
}

static void sched_save(const sched_entry *e)
{
    const char *path = getenv("STARSH_HIP_SCHEDULE");
    FILE *f = path ? fopen(path, "a") : NULL;
    if(f == NULL)
        return;
    fprintf(f, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g %lld %d %d %d "
            "%.17g %d %d %d %d %d %d %d %d %llu %d ", e->kd.family,
            e->kd.ndim, e->kd.beta, e->kd.nu, e->kd.noise, e->kd.sigma,
            e->kd.wave_k, e->kd.diag, e->n, e->nb, e->maxrank, e->os, e->tol,
            e->form, e->first, e->ladder, e->panel, e->gemm_bn[0],
            e->gemm_rw[0], e->gemm_bn[1], e->gemm_rw[1], e->geom,
            e->tiles ? e->ntile : 0);
    if(e->tiles && e->ntile > 0)
        for(int i = 0; i < e->ntile; i++)
            fputc(e->tiles[i] == -2 ? 'd' : '0'+e->tiles[i], f);
    else
        fputc('-', f);
    fputc('\n', f);
    fclose(f);
}

static const sched_entry *sched_find(const sched_entry *key)
{
    sched_load();
    for(int i = 0; i < nsched; i++)
        if(sched_same(&sched[i], key))
            return &sched[i];
    return NULL;
}

/* Per-tile schedule for a problem: the one recorded for the same signature,
 * else one recorded for the same geometry with other kernel parameters, as
 * in a parameter sweep over one point set.  A schedule only moves work
 * between buckets and stores tiles densely (exactly), so one recorded for
 * other parameters can cost time, never accuracy.  NULL when none applies
 * to ntile far-field tiles. */
static const sched_entry *sched_tiles(const sched_entry *key,
        STARSH_int ntile)
{
    const sched_entry *e = sched_find(key);
    if(e && e->tiles && e->ntile == ntile)
        return e;
    for(int i = nsched-1; i >= 0; i--)
        if(sched[i].tiles && sched[i].ntile == ntile &&
                sched_same_geom(&sched[i], key))
            return &sched[i];
    return NULL;
}

void hip_sched_clear(void)
{
    for(int i = 0; i < nsched; i++)
        free(sched[i].tiles);
    nsched = 0;
    sched_loaded = 1;
}

/* ---- state of a tuning run ---- */

/* Knobs of the trial being timed (instead of the table's) while tuning. */
static int trial_on = 0;
static sched_entry trial;
static const signed char **trial_tiles = NULL;  /* per problem, or NULL */
/* Outcome of every tile, per problem, recorded during a tuning run. */
static signed char **record_tiles = NULL;
/* Tiles still pending after the first bucket of the last mixed call. */
static long long pending_after_first = 0;

void hip_sched_apply(int mixed, const hip_batch *B)
{
    const sched_entry *e = NULL;
    sched_entry key;
    if(trial_on)
        e = &trial;
    else if(B->nprob > 0)
    {
        sched_key(mixed ? 1 : 0, B->hp, B->F[0], B->maxrank[0], B->tol[0],
                &key);
        e = sched_find(&key);
    }
    hip_set_tuned_panel(e ? e->panel : 0);
    for(int j = 0; j < 2; j++)
        hip_set_tuned_gemm(j, e ? e->gemm_bn[j] : 0, e ? e->gemm_rw[j] : 0);
    hip_mixed_tuned_ladder(mixed && e ? e->ladder : 0, trial_on);
}

void hip_sched_plan(const hip_batch *B, int q, int ladder, int *first,
        const signed char **tiles)
{
    *tiles = NULL;
    if(trial_on)
    {
        *first = trial.first >= 0 ? trial.first : DEFAULT_FIRST;
        *tiles = trial_tiles ? trial_tiles[q] : NULL;
        return;
    }
    if(forced_bucket >= 0)
    {
        *first = forced_bucket;
        return;
    }
    sched_entry key;
    sched_key(1, B->hp+q, B->F[q], B->maxrank[q], B->tol[q], &key);
    const sched_entry *e = sched_find(&key);
    *first = e && e->first >= 0 ? e->first : DEFAULT_FIRST;
    /* a per-tile schedule names buckets of its own ladder: used on that
     * ladder only */
    const sched_entry *s = use_tile_schedule ?
        sched_tiles(&key, B->F[q]->nblocks_far) : NULL;
    if(s && s->ladder == ladder)
        *tiles = s->tiles;
}

void hip_sched_record(int q, STARSH_int tile, int outcome)
{
    if(record_tiles && record_tiles[q])
        record_tiles[q][tile] = (signed char)outcome;
}

void hip_sched_pending_after_first(long long n)
{
    pending_after_first = n;
}

/* ---- the tuner ---- */

/* The tile lists of a format, kept so that every trial starts from the
 * format the caller passed (a compression rewrites it when it finds false
 * far-field tiles). */
typedef struct
{
    STARSH_int nfar, nnear;
    STARSH_int *far, *near;
} format_lists;

static int lists_save(const STARSH_blrf *F, format_lists *L)
{
    L->nfar = F->nblocks_far;
    L->nnear = F->nblocks_near;
    L->far = malloc(sizeof(STARSH_int)*2*(L->nfar > 0 ? L->nfar : 1));
    L->near = malloc(sizeof(STARSH_int)*2*(L->nnear > 0 ? L->nnear : 1));
    if(!L->far || !L->near)
        return STARSH_MALLOC_ERROR;
    if(L->nfar > 0)
        memcpy(L->far, F->block_far, sizeof(STARSH_int)*2*L->nfar);
    if(L->nnear > 0)
        memcpy(L->near, F->block_near, sizeof(STARSH_int)*2*L->nnear);
    return STARSH_SUCCESS;
}

/* Rebuilds *F from the saved lists if a trial changed them. */
static int lists_restore(STARSH_blrf *F, const format_lists *L)
{
    if(F->nblocks_far == L->nfar && F->nblocks_near == L->nnear)
        return STARSH_SUCCESS;
    STARSH_int *far = malloc(sizeof(STARSH_int)*2*(L->nfar > 0 ? L->nfar :
                1));
    STARSH_int *near = malloc(sizeof(STARSH_int)*2*(L->nnear > 0 ?
                L->nnear : 1));
    if(!far || !near)
    {
        free(far);
        free(near);
        return STARSH_MALLOC_ERROR;
    }
    memcpy(far, L->far, sizeof(STARSH_int)*2*L->nfar);
    memcpy(near, L->near, sizeof(STARSH_int)*2*L->nnear);
    STARSH_blrf *F2;
    int info = starsh_blrf_new_from_coo(&F2, F->problem, F->symm,
            F->row_cluster, F->col_cluster, L->nfar, far, L->nnear, near,
            F->type);
    if(info != STARSH_SUCCESS)
        return info;
    if(L->nfar == 0)
        free(far);
    if(L->nnear == 0)
        free(near);
    STARSH_blrf tmp = *F;
    *F = *F2;
    *F2 = tmp;
    starsh_blrf_free(F2);
    return STARSH_SUCCESS;
}

typedef struct
{
    STARSH_blrf **format;
    int nprob, form, onfly;
    const int *maxrank;
    const double *tol;
    format_lists *lists;
    int ntrials;
} tuner;

/* One compression of the batch with the knobs in `trial`; device time in
 * *seconds; the formats are restored afterwards. */
static int tune_run(tuner *T, double *seconds)
{
    STARSH_blrm **M = calloc(T->nprob, sizeof(*M));
    if(M == NULL)
        return STARSH_MALLOC_ERROR;
    trial_on = 1;
    int info = hip_compress(T->form, T->nprob, M, T->format, T->maxrank,
            T->tol, T->onfly);
    trial_on = 0;
    T->ntrials++;
    if(info == STARSH_SUCCESS)
    {
        starsh_hip_stats st;
        starsh_hip_get_stats(&st);
        *seconds = st.device_seconds;
        for(int q = 0; q < T->nprob; q++)
            starsh_blrm_free(M[q]);
        for(int q = 0; info == STARSH_SUCCESS && q < T->nprob; q++)
            info = lists_restore(T->format[q], T->lists+q);
    }
    free(M);
    return info;
}

/* Times each value of one knob (written through *knob) and keeps the
 * fastest; *best_t is the time to beat and is updated. */
static int tune_knob(tuner *T, int *knob, const int *cand, int ncand,
        double *best_t)
{
    int best = *knob;
    for(int i = 0; i < ncand; i++)
    {
        if(cand[i] == best)
            continue;
        *knob = cand[i];
        double t;
        int info = tune_run(T, &t);
        if(info != STARSH_SUCCESS)
        {
            *knob = best;
            return info;
        }
        if(t < *best_t)
        {
            *best_t = t;
            best = cand[i];
        }
    }
    *knob = best;
    return STARSH_SUCCESS;
}

/* Both GEMM knobs of operation j over (bn, rw) pairs. */
static int tune_gemm(tuner *T, int j, const int (*cand)[2], int ncand,
        double *best_t)
{
    int bb = trial.gemm_bn[j], br = trial.gemm_rw[j];
    for(int i = 0; i < ncand; i++)
    {
        trial.gemm_bn[j] = cand[i][0];
        trial.gemm_rw[j] = cand[i][1];
        double t;
        int info = tune_run(T, &t);
        if(info != STARSH_SUCCESS)
        {
            trial.gemm_bn[j] = bb;
            trial.gemm_rw[j] = br;
            return info;
        }
        if(t < *best_t)
        {
            *best_t = t;
            bb = cand[i][0];
            br = cand[i][1];
        }
    }
    trial.gemm_bn[j] = bb;
    trial.gemm_rw[j] = br;
    return STARSH_SUCCESS;
}

static int tune_mixed(tuner *T, starsh_hip_tuning *out, signed char ***rec)
{
    double t, best_t = 1e300;
    int info = STARSH_SUCCESS;
    for(int b = 0; b < NBUCKETS; b++)
    {
        trial.first = b;
        info = tune_run(T, &t);
        if(info != STARSH_SUCCESS)
            return info;
        out->first_seconds[b] = t;
        if(b == DEFAULT_FIRST)
            out->untuned_seconds = t;
        if(t < best_t)
        {
            best_t = t;
            out->first_bucket = b;
        }
        if(pending_after_first == 0 && b >= DEFAULT_FIRST)
            break;
    }
    if(out->untuned_seconds < 0)
    {
        trial.first = DEFAULT_FIRST;
        info = tune_run(T, &out->untuned_seconds);
        if(info != STARSH_SUCCESS)
            return info;
    }
    trial.first = out->first_bucket;
    /* ladder 1 (a 384-wide bucket, next bucket from the computed rank):
     * the same first width (the last bucket moves up one index) */
    trial.ladder = 1;
    if(trial.first == NBUCKETS-1)
        trial.first = NBUCKETS;
    info = tune_run(T, &out->ladder_seconds);
    if(info != STARSH_SUCCESS)
        return info;
    if(out->ladder_seconds < best_t)
    {
        best_t = out->ladder_seconds;
        out->ladder = 1;
        out->first_bucket = trial.first;
    }
    else
    {
        trial.ladder = 0;
        trial.first = out->first_bucket;
    }
    /* record every tile's outcome at the chosen first bucket */
    signed char **r = calloc(T->nprob, sizeof(*r));
    if(r == NULL)
        return STARSH_MALLOC_ERROR;
    *rec = r;
    for(int q = 0; q < T->nprob; q++)
    {
        const STARSH_int nf = T->format[q]->nblocks_far;
        r[q] = malloc(nf > 0 ? nf : 1);
        if(r[q] == NULL)
            return STARSH_MALLOC_ERROR;
        memset(r[q], -1, nf > 0 ? nf : 1);
    }
    record_tiles = r;
    info = tune_run(T, &t);
    record_tiles = NULL;
    if(info != STARSH_SUCCESS)
        return info;
    int whole = 1;
    for(int q = 0; q < T->nprob; q++)
        for(STARSH_int i = 0; whole && i < T->format[q]->nblocks_far; i++)
            whole = r[q][i] != -1;
    if(whole)
    {
        /* time the per-tile schedule; keep it only if it pays */
        trial_tiles = (const signed char **)r;
        info = tune_run(T, &out->schedule_seconds);
        trial_tiles = NULL;
        if(info != STARSH_SUCCESS)
            return info;
        if(out->schedule_seconds < best_t)
        {
            best_t = out->schedule_seconds;
            out->tile_schedule = 1;
        }
    }
    if(out->tile_schedule)
        trial_tiles = (const signed char **)r;
    static const int panels[2] = {4, 8};
    info = tune_knob(T, &trial.panel, panels, 2, &best_t);
    trial_tiles = NULL;
    out->best_seconds = best_t;
    return info;
}

static int tune_plain(tuner *T, starsh_hip_tuning *out)
{
    double best_t;
    int info = tune_run(T, &best_t);
    if(info != STARSH_SUCCESS)
        return info;
    out->untuned_seconds = best_t;
    /* sketch: wide column blocks reuse each generated entry most */
    static const int sketch[4][2] = {{256, 1}, {256, 2}, {128, 1}, {128, 2}};
    static const int proj[4][2] = {{128, 2}, {128, 1}, {64, 1}, {64, 2}};
    info = tune_gemm(T, 0, sketch, 4, &best_t);
    if(info == STARSH_SUCCESS)
        info = tune_gemm(T, 1, proj, 4, &best_t);
    static const int panels[2] = {4, 8};
    if(info == STARSH_SUCCESS)
        info = tune_knob(T, &trial.panel, panels, 2, &best_t);
    out->best_seconds = best_t;
    return info;
}

int starsh_hip_autotune_batch(STARSH_blrf **format, int nprob,
        const int *maxrank, const double *tol, int onfly, int form,
        starsh_hip_tuning *out)
{
    if(nprob <= 0 || format == NULL || maxrank == NULL || tol == NULL ||
            out == NULL || (form != 0 && form != 1))
        return STARSH_WRONG_PARAMETER;
    for(int q = 0; q < nprob; q++)
        if(format[q] == NULL)
            return STARSH_WRONG_PARAMETER;
    memset(out, 0, sizeof(*out));
    out->form = form;
    out->first_bucket = -1;
    out->schedule_seconds = -1.0;
    out->untuned_seconds = -1.0;
    out->ladder_seconds = -1.0;
    for(int b = 0; b < NBUCKETS; b++)
        out->first_seconds[b] = -1.0;
    tuner T = {format, nprob, form, onfly, maxrank, tol, NULL, 0};
    T.lists = calloc(nprob, sizeof(*T.lists));
    sched_entry *e = calloc(nprob, sizeof(*e));
    signed char **rec = NULL;
    int info = T.lists && e ? STARSH_SUCCESS : STARSH_MALLOC_ERROR;
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
        info = lists_save(format[q], T.lists+q);
    memset(&trial, 0, sizeof(trial));
    trial.form = form;
    trial.first = DEFAULT_FIRST;
    double t;
    if(info == STARSH_SUCCESS)
        info = tune_run(&T, &t);    /* warm-up, untimed */
    if(info == STARSH_SUCCESS)
        info = form ? tune_mixed(&T, out, &rec) : tune_plain(&T, out);
    out->qr_panel = trial.panel;
    out->sketch_bn = trial.gemm_bn[0];
    out->sketch_rw = trial.gemm_rw[0];
    out->proj_bn = trial.gemm_bn[1];
    out->proj_rw = trial.gemm_rw[1];
    out->ntrials = T.ntrials;
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
    {
        hip_problem hp;
        info = hip_driver_prepare(format[q], maxrank[q], tol[q], onfly, &hp);
        if(info != STARSH_SUCCESS)
            break;
        sched_key(form, &hp, format[q], maxrank[q], tol[q], e+q);
        hip_driver_release(&hp);
        e[q].first = form ? out->first_bucket : -1;
        e[q].ladder = form ? out->ladder : 0;
        e[q].panel = trial.panel;
        for(int j = 0; j < 2; j++)
        {
            e[q].gemm_bn[j] = trial.gemm_bn[j];
            e[q].gemm_rw[j] = trial.gemm_rw[j];
        }
        if(out->tile_schedule && format[q]->nblocks_far > 0 &&
                format[q]->nblocks_far < (1 << 28))
        {
            e[q].ntile = (int)format[q]->nblocks_far;
            e[q].tiles = rec[q];
            rec[q] = NULL;
        }
    }
    if(info == STARSH_SUCCESS)
    {
        sched_load();
        for(int q = 0; q < nprob; q++)
        {
            sched_save(e+q);
            sched_put(e+q);
        }
    }
    memset(&trial, 0, sizeof(trial));
    for(int q = 0; T.lists && q < nprob; q++)
    {
        free(T.lists[q].far);
        free(T.lists[q].near);
    }
    for(int q = 0; rec && q < nprob; q++)
        free(rec[q]);
    for(int q = 0; e && q < nprob; q++)
        free(e[q].tiles);
    free(rec);
    free(T.lists);
    free(e);
    return info;
}

int starsh_hip_autotune(STARSH_blrf *format, int maxrank, double tol,
        int onfly, int *first_out, double *seconds_out)
{
    if(format == NULL)
        return STARSH_WRONG_PARAMETER;
    starsh_hip_tuning r;
    int info = starsh_hip_autotune_batch(&format, 1, &maxrank, &tol, onfly,
            1, &r);
    if(info != STARSH_SUCCESS)
        return info;
    if(first_out)
        *first_out = r.first_bucket;
    for(int b = 0; seconds_out && b < NBUCKETS; b++)
        seconds_out[b] = r.first_seconds[b];
    return STARSH_SUCCESS;
}
