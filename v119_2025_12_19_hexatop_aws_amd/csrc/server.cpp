// Device-resident job server: topology-optimisation jobs of mixed mesh sizes
// admitted onto one GPU by device-memory footprint, run concurrently.
//
// The reference has no serving mode; its closest relative is the VRAM guard
// (R/src/Utils/Helpers.jl:177-244), which refuses a solve whose estimated
// footprint exceeds free memory.  Here every job is a handle with its own
// HIP stream; worker threads take jobs first-fit from a queue whenever a
// slot is free and the job's estimated bytes fit in the remaining budget, so
// jobs that arrive while others run start as soon as memory and a slot free
// up, and their kernels overlap on the device.  Every kernel of a job is
// deterministic and independent of co-scheduled work, so a job's result is
// bitwise the same alone and in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "handle.h"
#include "hexa_internal.h"

using namespace hexa;

namespace {

uint64_t splitmix(uint64_t seed, uint64_t index) {
    uint64_t z = seed + (index + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double unif(uint64_t seed, uint64_t index) { return (double)(splitmix(seed, index) >> 11) * (1.0 / 9007199254740992.0); }

struct Job {
    int64_t id = 0;
    hexa_job desc{};
    hexa_job_result res{};
    int64_t est = 0;
    bool done = false;
    bool claimed = false;    // a hexa_server_wait is waiting for this job
    std::vector<double> x;   // final design
    std::string err;         // the worker's error message when status != OK
};

}  // namespace

struct hexa_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::map<int64_t, Job*> jobs;
    std::vector<std::thread> workers;
    int64_t budget = 0, in_use = 0, next_id = 1;
    int running = 0;
    bool stop = false;
    double t0 = 0.0;
};

extern "C" int64_t hexa_estimate_bytes(const hexa_grid* g, const hexa_options* o) {
    if (!g) return 0;
    const double nn = (double)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    const double ne = (double)g->nx * g->ny * g->nz;
    const bool mixed = o && o->precision == HEXA_PREC_MIXED;
    // FP64 solver vectors (x r p Ap z minv best b diag guess) + mask/f/fixed
    double b = nn * (10 * 24 + 1 + 24 + 3);
    // multigrid hierarchy (level 0 plus ~1/7 for the coarser levels)
    const double lvl = mixed ? (4 * 12 + 6 + 24) : (4 * 24 + 24 + 24);
    b += nn * lvl * 8.0 / 7.0 + ne * (8 + 4 + 8) * 8.0 / 7.0;
    // filter nodal vectors and element work arrays
    b += nn * 6 * 8 + ne * 12 * 8;
    // the coarsest level's dense inverse and its factorisation panels, and
    // small buffers
    hexa_options d;
    if (!o) hexa_default_options(&d);
    int64_t nc = 0;
    hexa::mg_level_count(g->nx, g->ny, g->nz, o ? *o : d, &nc);
    b += (double)nc * nc * 8 + 3.0 * hexa::GJ_PANEL * nc * 8 + (64 << 20);
    return (int64_t)b;
}

namespace {

double wall_ms(const hexa_server* s) { return now_ms() - s->t0; }

// Builds the job's problem: support on one face (fixed, 3 nn flags), 1-4
// unit point loads at seeded nodes of the opposite face with seeded
// directions, as (DOF, value) pairs of the load vector.
void make_problem(const hexa_job& d, std::vector<uint8_t>& fixed, std::vector<int64_t>& fidx,
                  std::vector<double>& fval) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1, nn = NX * NY * NZ;
    fixed.assign(3 * nn, 0);
    const int face = ((d.support_face % 6) + 6) % 6;
    const int axis = face / 2;
    const bool hi = face % 2;
    const int64_t n[3] = {NX - 1, NY - 1, NZ - 1};
    const int ax1 = (axis + 1) % 3, ax2 = (axis + 2) % 3;
    // the nodes with coordinate axis == face value
    for (int64_t q = 0; q <= n[ax2]; ++q)
        for (int64_t p = 0; p <= n[ax1]; ++p) {
            int64_t c[3];
            c[axis] = hi ? n[axis] : 0;
            c[ax1] = p;
            c[ax2] = q;
            const int64_t id = c[0] + NX * (c[1] + NY * c[2]);
            fixed[3 * id] = fixed[3 * id + 1] = fixed[3 * id + 2] = 1;
        }
    std::map<int64_t, double> f;
    const int nl = std::max(1, std::min(4, d.n_loads));
    for (int l = 0; l < nl; ++l) {
        const double a = unif(d.seed, 4 * l), b = unif(d.seed, 4 * l + 1), z = 2.0 * unif(d.seed, 4 * l + 2) - 1.0,
                     phi = 2.0 * M_PI * unif(d.seed, 4 * l + 3);
        int64_t c[3];
        c[axis] = hi ? 0 : n[axis];   // opposite face
        c[ax1] = std::min<int64_t>(n[ax1], (int64_t)(a * (n[ax1] + 1)));
        c[ax2] = std::min<int64_t>(n[ax2], (int64_t)(b * (n[ax2] + 1)));
        const int64_t id = c[0] + NX * (c[1] + NY * c[2]);
        const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        f[3 * id] += s * std::cos(phi);
        f[3 * id + 1] += s * std::sin(phi);
        f[3 * id + 2] += z;
    }
    fidx.clear();
    fval.clear();
    for (const auto& kv : f) {
        fidx.push_back(kv.first);
        fval.push_back(kv.second);
    }
}

int32_t run_job(Job* jb) {
    const hexa_job& d = jb->desc;
    hexa_grid g{d.nx, d.ny, d.nz, 1.0, 1.0, 1.0};
    hexa_options o;
    hexa_default_options(&o);
    o.E0 = 1.0;
    o.Emin = 1e-9;
    o.penal = 3.0;
    o.precision = d.precision;
    hexa_handle* h = nullptr;
    int32_t st = hexa_create(&g, &o, &h);
    if (st != HEXA_OK) return st;
    std::vector<uint8_t> fixed;
    std::vector<int64_t> fidx;
    std::vector<double> fval;
    make_problem(d, fixed, fidx, fval);
    st = set_problem_impl(h, fixed.data(), nullptr, (int64_t)fidx.size(), fidx.data(), fval.data(), nullptr);
    const int64_t ne = (int64_t)d.nx * d.ny * d.nz;
    hexa_oc_params p;
    std::memset(&p, 0, sizeof p);
    p.volfrac = d.volfrac;
    p.move = d.move > 0 ? d.move : 0.2;
    p.rmin = d.rmin > 0 ? d.rmin : 1.5;
    p.filter_tol = 1e-10;
    p.filter_max_iter = 500;
    p.bisect_tol = 1e-12;
    hexa_default_solve_params(&p.solve);
    p.solve.tol = d.tol > 0 ? d.tol : 1e-8;
    p.solve.precond = HEXA_PRECOND_MG;
    p.solve.semantics = HEXA_SEMANTICS_CPU;
    p.solve.max_iter = 2000;
    // the design stays on the device (the handle's OC design array) from the
    // uniform start to the last iteration, then comes back once
    if (st == HEXA_OK) st = oc_check(h, &p);
    if (st == HEXA_OK) {
        launch_fill<double>(h->stream, ne, d.volfrac, h->d_e0);
        st = check_launch();
    }
    int solves = 0, iters = 0;
    hexa_oc_report rep{};
    for (int it = 0; it < d.oc_iters && st == HEXA_OK; ++it) {
        p.solve.warm_start = it > 0;
        st = oc_step_core(h, &p, &rep, now_ms());
        if (st == HEXA_E_NOT_CONVERGED) st = HEXA_OK;
        ++solves;
        iters += rep.solve_iterations;
    }
    jb->x.assign(ne, 0.0);
    if (st == HEXA_OK &&
        (hipMemcpyAsync(jb->x.data(), h->d_e0, sizeof(double) * ne, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess)) {
        set_error("design download failed");
        st = HEXA_E_DEVICE;
    }
    jb->res.oc_iters_done = solves;
    jb->res.solve_iterations = iters;
    jb->res.compliance = rep.compliance;
    jb->res.volume = rep.volume;
    jb->res.device_bytes = h->bytes_allocated;
    hexa_destroy(h);
    return st;
}

void worker(hexa_server* s) {
    for (;;) {
        Job* jb = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop && s->queue.empty()) return;
                // first fit: the oldest queued job whose footprint fits
                for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
                    if ((*it)->est <= s->budget - s->in_use || s->running == 0) {
                        jb = *it;
                        s->queue.erase(it);
                        break;
                    }
                }
                if (jb) break;
                s->cv_work.wait(lk);
            }
            s->in_use += jb->est;
            s->running += 1;
            jb->res.t_start_ms = wall_ms(s);
        }
        const int32_t st = run_job(jb);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            jb->res.status = st;
            if (st != HEXA_OK) jb->err = hexa_last_error();
            jb->res.t_end_ms = wall_ms(s);
            jb->done = true;
            s->in_use -= jb->est;
            s->running -= 1;
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int32_t hexa_server_create(int32_t workers, int64_t mem_budget_bytes, hexa_server** out) {
    if (!out || workers < 1 || workers > 64 || mem_budget_bytes <= 0) {
        set_error("invalid server arguments");
        return HEXA_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device"); return HEXA_E_DEVICE; }
    hexa_server* s = new hexa_server();
    s->budget = mem_budget_bytes;
    s->t0 = now_ms();
    for (int i = 0; i < workers; ++i) s->workers.emplace_back(worker, s);
    *out = s;
    return HEXA_OK;
}

int32_t hexa_server_submit(hexa_server* s, const hexa_job* job, int64_t* id) {
    if (!s || !job || !id) { set_error("null argument"); return HEXA_E_INVALID; }
    if (job->nx < 2 || job->ny < 2 || job->nz < 2 || job->oc_iters < 1 || !(job->volfrac > 0 && job->volfrac < 1)) {
        set_error("invalid job");
        return HEXA_E_INVALID;
    }
    Job* jb = new Job();
    jb->desc = *job;
    hexa_grid g{job->nx, job->ny, job->nz, 1.0, 1.0, 1.0};
    hexa_options o;
    hexa_default_options(&o);
    o.precision = job->precision;
    jb->est = hexa_estimate_bytes(&g, &o);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) { delete jb; set_error("server is shutting down"); return HEXA_E_INVALID; }
        jb->id = s->next_id++;
        jb->res.id = jb->id;
        jb->res.t_submit_ms = wall_ms(s);
        jb->res.est_bytes = jb->est;
        s->jobs[jb->id] = jb;
        s->queue.push_back(jb);
        *id = jb->id;
    }
    s->cv_work.notify_all();
    return HEXA_OK;
}

int32_t hexa_server_wait(hexa_server* s, int64_t id, hexa_job_result* res, double* x_out) {
    if (!s) { set_error("null server"); return HEXA_E_INVALID; }
    Job* jb = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) { set_error("unknown job id"); return HEXA_E_INVALID; }
        jb = it->second;
        // one waiter per job: it alone erases and frees the entry
        if (jb->claimed) { set_error("job %lld is already being waited for", (long long)id); return HEXA_E_INVALID; }
        jb->claimed = true;
        s->cv_done.wait(lk, [&] { return jb->done; });
        s->jobs.erase(id);
    }
    if (res) *res = jb->res;
    if (x_out) std::memcpy(x_out, jb->x.data(), sizeof(double) * jb->x.size());
    const int32_t st = jb->res.status;
    // error messages are per thread: hand the worker's to the waiter
    if (st != HEXA_OK) set_error("job %lld: %s", (long long)id, jb->err.c_str());
    delete jb;
    return st;
}

double hexa_server_clock_ms(hexa_server* s) { return s ? wall_ms(s) : 0.0; }

int32_t hexa_server_destroy(hexa_server* s) {
    if (!s) return HEXA_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->jobs) delete kv.second;
    delete s;
    return HEXA_OK;
}

}  // extern "C"
