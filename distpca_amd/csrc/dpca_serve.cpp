// Projection of new samples onto resident loadings, and the serving loop:
// continuous batching of projection requests into one grouped launch per
// batch (replayed from a captured HIP graph) next to cohort-PCA jobs that are
// admitted by their HBM footprint and run concurrently on their own streams.
//
// No reference counterpart: DistPCA computes neither loadings nor
// projections (its path ends at the left singular vectors,
// src/methods.cpp:1072-1118).  Scores of a new sample are u = z W with
// W = Z U diag(1/sigma^2) and z standardised with the training per-SNP
// table, so projecting a training sample reproduces its row of U.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "dpca.h"
#include "dpca_internal.h"
#include "dpca_model.h"

namespace dpca {

static double mono_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// The SNP split of the grouped product depends only on the model, never on
// the batch width: every sample's partial sums then cover the same SNP ranges
// in the same order whatever batch it lands in, so a request's scores are
// bitwise the same alone and inside any batch.
ProjPlan plan_projection(const dpca_model *m, int64_t ldb)
{
    ProjPlan pl;
    pl.ldb = ldb;
    const int64_t samples = 4 * ldb;
    if (m->precision == DPCA_FP64) {
        pl.splits = (int)std::max<int64_t>(1, (m->Npad + kProjSplitRowsF64 - 1) / kProjSplitRowsF64);
        pl.part_bytes = (size_t)pl.splits * samples * m->kp * sizeof(double);
    } else {
        pl.sa_tiles = zt_i8_tiles(m->k, m->ns, ldb, 8);
        pl.rps = proj_rps_mx(m->Npad);
        pl.nsplit = (int)((m->Npad + pl.rps - 1) / pl.rps);
        pl.part_bytes = (size_t)pl.nsplit * samples * m->kp * sizeof(float);
    }
    return pl;
}

hipError_t launch_projection(hipStream_t st, const dpca_model *m, const ProjPlan &pl,
                             const uint8_t *batch, void *part, double *scores)
{
    const int64_t n = 4 * pl.ldb * m->kp;
    hipError_t e;
    if (m->precision == DPCA_FP64) {
        e = launch_zt_f64(st, batch, m->Npad, pl.ldb, m->zval, m->W, m->k, m->kp, pl.splits,
                          (double *)part);
        if (e == hipSuccess) e = launch_reduce_parts(st, (const double *)part, pl.splits, n, scores, 0);
    } else {
        e = launch_zt_i8(st, batch, nullptr, m->Npad, pl.ldb, m->UW8, m->smax, m->k, m->ns, pl.sa_tiles, 8,
                         pl.rps, pl.nsplit, (float *)part);
        if (e == hipSuccess) e = launch_reduce_f32(st, (const float *)part, pl.nsplit, n, scores, 0);
    }
    return e;
}

}  // namespace dpca

using namespace dpca;

#define SCHECK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));     \
            return -2;                                                                        \
        }                                                                                     \
    } while (0)

// ------------------------------------------------------------- the server --
struct Ticket {
    int kind = 0;                    // DPCA_TICKET_PROJECT / DPCA_TICKET_JOB
    int status = 0;                  // 0 queued, 1 running, 2 done, <0 error
    // projection
    const uint8_t *payload = nullptr;
    int n = 0;
    // job
    dpca_gen_spec spec{};
    dpca_params params{};
    long long footprint = 0;
    // results
    std::vector<double> out;
    double t_submit = 0, t_start = 0, t_done = 0;
    long long batch = -1;
    int batch_samples = 0;
    char err[256] = "";
};

struct GraphClass {
    ProjPlan plan;
    hipGraphExec_t exec = nullptr;
};

struct dpca_server {
    dpca_server_params p{};
    dpca_model *model = nullptr;
    double t0 = 0;
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    long long next_ticket = 1;
    std::map<long long, Ticket> tickets;
    std::deque<long long> proj_q, job_q;
    long long budget_free = 0;
    long long job_capacity = 0;      // budget left for jobs after the model's share
    int jobs_running = 0;
    // projection resources (device)
    hipStream_t pst = nullptr;
    uint8_t *stage = nullptr, *batch = nullptr;
    void *part = nullptr;
    double *scores = nullptr;
    int64_t cap_bytes = 0;           // bytes per SNP row of the batch matrix
    std::vector<GraphClass> classes;
    std::vector<double> host_scores;
    std::thread batcher;
    std::vector<std::thread> workers;
    // stats
    long long batches = 0, batch_samples_total = 0, jobs_done = 0, proj_done = 0;
    std::atomic<long long> rejected{0};
};

// Continuous batching: everything queued when the previous batch finishes
// (FIFO, up to the batch capacity) forms the next batch.
static void batcher_loop(dpca_server *s)
{
    (void)hipSetDevice(s->p.device);
    const dpca_model *m = s->model;
    const int kp = m->kp;
    for (long long bid = 0;; bid++) {
        std::vector<long long> ids;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->proj_q.empty(); });
            if (s->stop && s->proj_q.empty()) return;
            int64_t used = 0;
            while (!s->proj_q.empty()) {
                Ticket &t = s->tickets[s->proj_q.front()];
                const int64_t nb = (t.n + 3) / 4;
                if (used + nb > s->cap_bytes && !ids.empty()) break;
                used += nb;
                ids.push_back(s->proj_q.front());
                s->proj_q.pop_front();
                t.status = 1;
                t.t_start = mono_s() - s->t0;
            }
        }
        // byte offsets of each request in the batch, and the graph class
        std::vector<int64_t> off(ids.size());
        int64_t used = 0;
        std::vector<const uint8_t *> pay(ids.size());
        std::vector<int> ns(ids.size());
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < ids.size(); i++) {
                Ticket &t = s->tickets[ids[i]];
                off[i] = used;
                used += (t.n + 3) / 4;
                pay[i] = t.payload;
                ns[i] = t.n;
            }
        }
        size_t ci = 0;
        while (ci + 1 < s->classes.size() && s->classes[ci].plan.ldb < used) ci++;
        const GraphClass &gc = s->classes[ci];
        const int64_t ldb = gc.plan.ldb;
        hipError_t e = hipSuccess;
        int64_t soff = 0;
        for (size_t i = 0; i < ids.size() && e == hipSuccess; i++) {
            const int64_t nb = (ns[i] + 3) / 4;
            e = hipMemcpyAsync(s->stage + soff, pay[i], (size_t)(m->N * nb), hipMemcpyHostToDevice,
                               s->pst);
            if (e == hipSuccess)
                e = launch_scatter_cols(s->pst, s->stage + soff, m->N, nb, ns[i], s->batch, ldb,
                                        off[i]);
            soff += m->N * nb;
        }
        if (e == hipSuccess) e = launch_fill_cols(s->pst, s->batch, m->N, ldb, used, ldb);
        if (e == hipSuccess && m->Npad > m->N)
            e = launch_fill_cols(s->pst, s->batch + m->N * ldb, m->Npad - m->N, ldb, 0, ldb);
        if (e == hipSuccess) e = hipGraphLaunch(gc.exec, s->pst);
        const int64_t rows_used = 4 * used;
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->host_scores.data(), s->scores,
                               (size_t)(rows_used * kp * sizeof(double)), hipMemcpyDeviceToHost,
                               s->pst);
        if (e == hipSuccess) e = hipStreamSynchronize(s->pst);
        const double now = mono_s() - s->t0;
        std::lock_guard<std::mutex> lk(s->mu);
        s->batches++;
        s->batch_samples_total += rows_used;
        for (size_t i = 0; i < ids.size(); i++) {
            Ticket &t = s->tickets[ids[i]];
            t.batch = bid;
            t.batch_samples = (int)rows_used;
            t.t_done = now;
            if (e != hipSuccess) {
                t.status = -2;
                snprintf(t.err, sizeof t.err, "projection batch: %s", hipGetErrorString(e));
                continue;
            }
            t.out.resize((size_t)t.n * m->nsv);
            const int64_t r0 = 4 * off[i];
            for (int a = 0; a < t.n; a++)
                for (int j = 0; j < m->nsv; j++)
                    t.out[(size_t)a * m->nsv + j] = s->host_scores[(size_t)(r0 + a) * kp + j];
            t.status = 2;
            s->proj_done++;
        }
        s->cv.notify_all();
    }
}

// Cohort-PCA jobs: first queued job whose footprint fits the free HBM budget.
static void worker_loop(dpca_server *s)
{
    (void)hipSetDevice(s->p.device);
    for (;;) {
        long long id = 0;
        Ticket job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop && s->job_q.empty()) return;
                auto it = std::find_if(s->job_q.begin(), s->job_q.end(), [&](long long j) {
                    return s->tickets[j].footprint <= s->budget_free;
                });
                if (it != s->job_q.end()) {
                    id = *it;
                    s->job_q.erase(it);
                    break;
                }
                s->cv.wait(lk);
            }
            Ticket &t = s->tickets[id];
            s->budget_free -= t.footprint;
            s->jobs_running++;
            t.status = 1;
            t.t_start = mono_s() - s->t0;
            job = t;
        }
        dpca_ctx *c = nullptr;
        set_job_contexts(true);
        int r = dpca_open_synthetic(&job.spec, &job.params, &c);
        set_job_contexts(false);
        std::vector<double> sigma;
        if (!r) {
            // measured at 20 requests/s: graph-replayed job iterations left
            // the jobs' latency as it was and raised the projections' p99
            // by ~30%, so jobs enqueue their iterations launch by launch
            ctx_eager_iterations(c);
            int M, nsv, nrhs, bsize;
            long long N;
            dpca_get_dims(c, &M, &N, &nsv, &nrhs, &bsize);
            sigma.resize(nsv);
            dpca_report rep;
            r = dpca_run(c, nullptr, sigma.data(), &rep);
            dpca_close(c);
        }
        const char *err = r ? dpca_last_error() : nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            Ticket &t = s->tickets[id];
            s->budget_free += t.footprint;
            s->jobs_running--;
            t.out = sigma;
            t.status = r ? r : 2;
            t.t_done = mono_s() - s->t0;
            if (err) snprintf(t.err, sizeof t.err, "%s", err);
            if (!r) s->jobs_done++;
        }
        s->cv.notify_all();
    }
}

extern "C" {

int dpca_model_info(dpca_model *m, long long *N, int *nsv, long long *bytes)
{
    if (!m) return -1;
    if (N) *N = m->N;
    if (nsv) *nsv = m->nsv;
    if (bytes) *bytes = (long long)m->bytes;
    return 0;
}

void dpca_model_free(dpca_model *m)
{
    if (!m) return;
    (void)hipSetDevice(m->dev);
    void *ptrs[] = {m->zval, m->mu, m->dsc, m->W, m->UW8, m->smax};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
}

int dpca_model_project(dpca_model *m, const uint8_t *packed, int n, double *scores)
{
    if (!m || !packed || !scores || n <= 0) {
        set_error("dpca_model_project: bad arguments");
        return -1;
    }
    SCHECK(hipSetDevice(m->dev));
    const int64_t nb = (n + 3) / 4;
    const int64_t ldb = round_up(nb, 64);
    const ProjPlan pl = plan_projection(m, ldb);
    uint8_t *stage = nullptr, *batch = nullptr;
    void *part = nullptr;
    double *sc = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&stage, (size_t)(m->N * nb));
    if (e == hipSuccess) e = hipMalloc((void **)&batch, (size_t)(m->Npad * ldb));
    if (e == hipSuccess) e = hipMalloc(&part, pl.part_bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&sc, (size_t)(4 * ldb * m->kp * sizeof(double)));
    if (e == hipSuccess) e = hipMemcpyAsync(stage, packed, (size_t)(m->N * nb), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_scatter_cols(st, stage, m->N, nb, n, batch, ldb, 0);
    if (e == hipSuccess) e = launch_fill_cols(st, batch, m->N, ldb, nb, ldb);
    if (e == hipSuccess && m->Npad > m->N)
        e = launch_fill_cols(st, batch + m->N * ldb, m->Npad - m->N, ldb, 0, ldb);
    if (e == hipSuccess) e = launch_projection(st, m, pl, batch, part, sc);
    std::vector<double> h;
    if (e == hipSuccess) {
        h.resize((size_t)n * m->kp);
        e = hipMemcpyAsync(h.data(), sc, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error("dpca_model_project: %s", hipGetErrorString(e));
        rc = -2;
    } else {
        for (int a = 0; a < n; a++)
            for (int j = 0; j < m->nsv; j++) scores[(size_t)a * m->nsv + j] = h[(size_t)a * m->kp + j];
    }
    void *ptrs[] = {stage, batch, part, sc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int dpca_server_params_init(dpca_server_params *p)
{
    if (!p) return -1;
    memset(p, 0, sizeof *p);
    p->device = 0;
    p->hbm_budget = 64ll << 30;
    p->batch_capacity = 8192;
    p->workers = 2;
    return 0;
}

int dpca_server_create(const dpca_server_params *p, dpca_model *model, dpca_server **out)
{
    if (!p || !out || p->workers < 0 || p->batch_capacity < 4) {
        set_error("dpca_server_create: bad arguments");
        return -1;
    }
    SCHECK(hipSetDevice(p->device));
    dpca_server *s = new dpca_server();
    s->p = *p;
    s->model = model;
    s->t0 = mono_s();
    s->budget_free = p->hbm_budget;
    auto fail = [&](int rc) {
        s->stop = true;
        dpca_server_destroy(s);
        return rc;
    };
    if (model) {
        // batch classes: 256, 1024, 4096 samples and the capacity; one
        // captured graph (grouped projection + split reduction) per class
        s->cap_bytes = round_up((p->batch_capacity + 3) / 4, 64);
        std::vector<int64_t> cls;
        for (int64_t b : {64ll, 256ll, 1024ll})
            if (b < s->cap_bytes) cls.push_back(b);
        cls.push_back(s->cap_bytes);
        size_t part_max = 0;
        for (int64_t b : cls) {
            GraphClass gc;
            gc.plan = plan_projection(model, b);
            part_max = std::max(part_max, gc.plan.part_bytes);
            s->classes.push_back(gc);
        }
        if (hipStreamCreateWithFlags(&s->pst, hipStreamNonBlocking) != hipSuccess ||
            hipMalloc((void **)&s->stage, (size_t)(model->N * s->cap_bytes)) != hipSuccess ||
            hipMalloc((void **)&s->batch, (size_t)(model->Npad * s->cap_bytes)) != hipSuccess ||
            hipMalloc(&s->part, part_max) != hipSuccess ||
            hipMalloc((void **)&s->scores, (size_t)(4 * s->cap_bytes * model->kp * sizeof(double))) !=
                hipSuccess) {
            set_error("dpca_server_create: device allocation failed");
            return fail(-3);
        }
        const long long reserved = (long long)(model->N * s->cap_bytes + model->Npad * s->cap_bytes +
                                               part_max + 4 * s->cap_bytes * model->kp * 8);
        s->budget_free -= reserved + (long long)model->bytes;
        s->host_scores.resize((size_t)(4 * s->cap_bytes * model->kp));
        for (auto &gc : s->classes) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(s->pst, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) e = launch_projection(s->pst, model, gc.plan, s->batch, s->part, s->scores);
            hipError_t e2 = hipStreamEndCapture(s->pst, &g);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                set_error("dpca_server_create: graph capture: %s", hipGetErrorString(e));
                return fail(-2);
            }
        }
        s->batcher = std::thread(batcher_loop, s);
    }
    s->job_capacity = s->budget_free;
    for (int w = 0; w < p->workers; w++) s->workers.emplace_back(worker_loop, s);
    *out = s;
    return 0;
}

int dpca_server_submit_project(dpca_server *s, const uint8_t *packed, int n, long long *ticket)
{
    if (!s || !packed || n <= 0 || !ticket || !s->model) {
        set_error("dpca_server_submit_project: bad arguments (or no model)");
        return -1;
    }
    if ((n + 3) / 4 > s->cap_bytes) {
        set_error("dpca_server_submit_project: %d samples exceed the batch capacity", n);
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const long long id = s->next_ticket++;
    Ticket &t = s->tickets[id];
    t.kind = DPCA_TICKET_PROJECT;
    t.payload = packed;
    t.n = n;
    t.t_submit = mono_s() - s->t0;
    s->proj_q.push_back(id);
    *ticket = id;
    s->cv.notify_all();
    return 0;
}

int dpca_server_submit_job(dpca_server *s, const dpca_gen_spec *spec, const dpca_params *p,
                           long long *ticket)
{
    if (!s || !spec || !p || !ticket) {
        set_error("dpca_server_submit_job: bad arguments");
        return -1;
    }
    if (s->p.workers < 1) {
        set_error("dpca_server_submit_job: the server has no job workers");
        return -1;
    }
    dpca_params q = *p;
    q.device = s->p.device;
    long long fp = 0;
    set_job_contexts(true);
    int r = dpca_footprint(spec->M, spec->N, &q, &fp);
    set_job_contexts(false);
    if (r) return r;
    fp += (long long)spec->N * (long long)std::max(1, spec->npop) * 4 + (64ll << 20);  // generator tables, slack
    std::lock_guard<std::mutex> lk(s->mu);
    if (fp > s->job_capacity) {  // could never be admitted
        s->rejected++;
        set_error("job footprint %lld bytes exceeds the HBM budget for jobs (%lld)", fp,
                  s->job_capacity);
        return -5;
    }
    const long long id = s->next_ticket++;
    Ticket &t = s->tickets[id];
    t.kind = DPCA_TICKET_JOB;
    t.spec = *spec;
    t.params = q;
    t.footprint = fp;
    t.t_submit = mono_s() - s->t0;
    s->job_q.push_back(id);
    *ticket = id;
    s->cv.notify_all();
    return 0;
}

int dpca_server_wait(dpca_server *s, long long ticket, double *out, long long cap,
                     dpca_ticket_info *info)
{
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) {
        set_error("dpca_server_wait: unknown ticket %lld", ticket);
        return -1;
    }
    s->cv.wait(lk, [&] { return it->second.status >= 2 || it->second.status < 0; });
    Ticket &t = it->second;
    if (info) {
        memset(info, 0, sizeof *info);
        info->kind = t.kind;
        info->status = t.status;
        info->n_out = (long long)t.out.size();
        info->t_submit = t.t_submit;
        info->t_start = t.t_start;
        info->t_done = t.t_done;
        info->batch = t.batch;
        info->batch_samples = t.batch_samples;
        info->footprint = t.footprint;
    }
    if (out)
        for (long long i = 0; i < std::min<long long>(cap, (long long)t.out.size()); i++)
            out[i] = t.out[i];
    const int st = t.status;
    if (st < 0) set_error("%s", t.err);
    s->tickets.erase(it);
    return st < 0 ? st : 0;
}

int dpca_server_stats(dpca_server *s, dpca_server_counters *st)
{
    if (!s || !st) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    st->batches = s->batches;
    st->batch_samples = s->batch_samples_total;
    st->projections_done = s->proj_done;
    st->jobs_done = s->jobs_done;
    st->jobs_rejected = s->rejected.load();
    st->budget_free = s->budget_free;
    return 0;
}

void *dpca_host_alloc(long long bytes)
{
    void *p = nullptr;
    if (bytes <= 0 || hipHostMalloc(&p, (size_t)bytes) != hipSuccess) {
        set_error("dpca_host_alloc(%lld) failed", bytes);
        return nullptr;
    }
    return p;
}

void dpca_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

void dpca_server_destroy(dpca_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->batcher.joinable()) s->batcher.join();
    for (auto &w : s->workers)
        if (w.joinable()) w.join();
    (void)hipSetDevice(s->p.device);
    for (auto &gc : s->classes)
        if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
    void *ptrs[] = {s->stage, s->batch, s->part, s->scores};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (s->pst) (void)hipStreamDestroy(s->pst);
    delete s;
}

}  // extern "C"
