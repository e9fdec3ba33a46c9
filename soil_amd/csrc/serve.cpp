// Served stream: simulation jobs (one box each) arrive at any time, are
// admitted against an HBM budget and a slot limit, stay resident on the
// device, and step together in the batched form (one launch per phase for all
// admitted boxes). Composition changes at chunk boundaries: a job that reaches
// its step count freezes on device (its step limit) and is retired at the end
// of the chunk; queued jobs that fit join the batch there. New boxes are set
// up (allocated and uploaded) on the host thread while the current chunk runs
// on the device.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "box.hpp"

using namespace soil_hip;

#define HIPCHK(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return set_error(SOIL_HIP_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int id = 0;
    soil_hip_box_desc desc{};
    std::vector<double> pos, vel, mass, rad;
    int steps = 0;
    double k = 1000.0;
    bool keep_result = false;
    double t_arrive = 0, t_admit = 0, t_done = 0;
    double t_last = 0;  // the job's last state update (admission, then each chunk's completion)
    soil_hip_box* box = nullptr;
    size_t bytes = 0;
    int done_steps = 0;  // steps this job has taken on the device
    bool failed = false;
    std::vector<double> out_pos, out_vel;
    size_t n = 0;
    // restart files: the job's input (when submitted from one), its first step
    // number, and where an eviction request wants its state written
    std::string restart_in;
    int restart_in_format = 0;
    uint64_t start_step = 0;
    std::string evict_path;
    int evict_format = 0;
    bool evicted = false;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double idx = p * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(idx), hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

}  // namespace

struct soil_hip_server {
    soil_hip_ctx* ctx = nullptr;
    size_t budget = 0;
    int max_slots = 256;
    int chunk = 16;
    bool use_graph = true;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Job>> pending;
    std::vector<std::unique_ptr<Job>> active;   // in the batch (or ready to join)
    std::vector<std::unique_ptr<Job>> done;
    soil_hip_batch* batch = nullptr;
    uint64_t batch_steps = 0;                   // steps run by the current batch
    uint64_t server_steps = 0;
    size_t used = 0;
    int next_id = 0;
    std::thread worker;
    std::atomic<bool> running{false};
    std::atomic<bool> draining{false};
    int error = 0;
    std::string error_msg;
    // finished jobs' RestartData (empty dir: none written)
    std::string out_dir;
    int out_format = 0;
    // eviction requests for jobs on the device: id -> (restart path, format)
    std::map<int, std::pair<std::string, int>> evict_req;
    // statistics
    // one sample per (job, step): the wall time between the job's consecutive
    // state updates (chunk completions, from its admission on) over the steps
    // each update brought, gaps for re-batching, admission and read-back included
    std::vector<double> step_ms;
    std::vector<double> gpu_step_ms;            // the chunk's device time over its steps, per (job, step)
    std::vector<double> gap_ms;                 // wall time between a job's consecutive updates
    double particle_steps = 0.0;
    double busy_s = 0.0;
    uint64_t reconfigs = 0;
    double reconfig_s = 0.0;
    double active_sum = 0.0;
    uint64_t chunks = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

int make_job_box(soil_hip_server* sv, Job* j) {
    soil_hip_box* b = nullptr;
    int rc = soil_hip_box_create(sv->ctx, &j->desc, &b);
    if (rc) return rc;
    if (!j->restart_in.empty()) {
        rc = soil_hip_box_read_restart(b, j->restart_in.c_str(), j->restart_in_format, &j->start_step);
        j->n = b->nlocal;
    } else {
        rc = soil_hip_upload(b, j->n, j->pos.data(), j->vel.data(), nullptr, j->mass.data(), j->rad.data(), nullptr,
                             nullptr, nullptr);
    }
    if (!rc) rc = soil_hip_set_consumer(b, 1, j->k);
    if (rc) {
        soil_hip_box_destroy(b);
        return rc;
    }
    j->bytes = box_bytes_estimate(b);
    j->box = b;
    return 0;
}

// Finishes jobs whose step count is reached: copies out results, frees boxes.
// The batch is released first (its membership changes and the boxes it points
// at are about to be destroyed); the loop builds the next one.
const char* restart_ext(int format) { return format == SOIL_HIP_RESTART_JSON ? "json" : "bin"; }

int retire(soil_hip_server* sv) {
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (auto& j : sv->active) {
            auto it = sv->evict_req.find(j->id);
            if (it == sv->evict_req.end()) continue;
            j->evict_path = it->second.first;
            j->evict_format = it->second.second;
            sv->evict_req.erase(it);
            j->evicted = j->done_steps < j->steps;  // a job that just finished is simply done
        }
    }
    bool any = false;
    for (auto& j : sv->active) any = any || j->failed || j->evicted || j->done_steps >= j->steps;
    if (!any) return 0;
    if (sv->batch) {
        batch_free(sv->batch, false);
        sv->batch = nullptr;
    }
    int err = 0;
    std::vector<std::unique_ptr<Job>> keep;
    for (auto& j : sv->active) {
        if (!j->failed && !j->evicted && j->done_steps < j->steps) {
            keep.push_back(std::move(j));
            continue;
        }
        j->t_done = sv->now();
        const uint64_t at = j->start_step + (uint64_t)j->done_steps;
        if (!j->failed && j->evicted) {
            // preempted: its state goes to the requested restart file (write_restart semantics)
            if (soil_hip_box_write_restart(j->box, j->evict_path.c_str(), j->evict_format, at)) j->failed = true;
        } else if (!j->failed && !sv->out_dir.empty()) {
            const std::string path = sv->out_dir + "/restart_" + std::to_string(at) + "_job" + std::to_string(j->id) +
                                     "." + restart_ext(sv->out_format);
            if (soil_hip_box_write_restart(j->box, path.c_str(), sv->out_format, at)) j->failed = true;
        }
        if (j->keep_result && !j->failed && !j->evicted) {
            uint64_t c[8];
            int rc = soil_hip_get_counts(j->box, c);
            const size_t na = rc ? 0 : c[0] + c[1];
            std::vector<double> p(3 * na), v(3 * na);
            if (!rc)
                rc = soil_hip_get_atoms(j->box, na, p.data(), v.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                        nullptr, nullptr);
            if (rc) {
                j->failed = true;
                err = rc;
            } else {
                j->n = (size_t)c[0];  // atoms that left through a non-periodic face are gone
                j->out_pos.assign(p.begin(), p.begin() + 3 * j->n);
                j->out_vel.assign(v.begin(), v.begin() + 3 * j->n);
            }
        }
        sv->used -= j->bytes;
        soil_hip_box_destroy(j->box);
        j->box = nullptr;
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->done.push_back(std::move(j));
    }
    sv->active.swap(keep);
    return err;
}

// Takes queued jobs that fit the budget and the slot table, in arrival order,
// and sets their boxes up on the device (allocation + upload on the box's own
// stream, so it overlaps a running chunk).
void admit(soil_hip_server* sv, std::vector<std::unique_ptr<Job>>& incoming) {
    while (true) {
        std::unique_ptr<Job> j;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (sv->pending.empty() || (int)(sv->active.size() + incoming.size()) >= sv->max_slots) return;
            Job* f = sv->pending.front().get();
            if (!f->box && make_job_box(sv, f)) {
                f->failed = true;
                f->t_done = sv->now();
                sv->done.push_back(std::move(sv->pending.front()));
                sv->pending.pop_front();
                continue;
            }
            if (sv->used + f->bytes > sv->budget) return;  // FIFO: wait for memory to free up
            sv->used += f->bytes;
            j = std::move(sv->pending.front());
            sv->pending.pop_front();
        }
        j->t_admit = sv->now();
        j->t_last = j->t_admit;
        int rc = soil_hip_setup(j->box);
        if (!rc) {
            const uint64_t limit = (uint64_t)j->steps;  // the box freezes on device after its own steps
            if (hipMemcpyAsync(&j->box->d_st->step_limit, &limit, 8, hipMemcpyHostToDevice, j->box->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(j->box->stream) != hipSuccess)
                rc = SOIL_HIP_EDEVICE;
        }
        if (rc) {
            j->failed = true;
            j->t_done = sv->now();
            sv->used -= j->bytes;
            soil_hip_box_destroy(j->box);
            j->box = nullptr;
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->done.push_back(std::move(j));
            continue;
        }
        incoming.push_back(std::move(j));
    }
}

int rebatch(soil_hip_server* sv) {
    auto t = Clock::now();
    if (sv->batch) {
        batch_free(sv->batch, false);
        sv->batch = nullptr;
    }
    std::vector<soil_hip_box*> boxes;
    for (auto& j : sv->active) boxes.push_back(j->box);
    if (!boxes.empty()) {
        int rc = batch_make(sv->ctx, boxes.data(), (int)boxes.size(), &sv->batch);
        if (rc) return rc;
    }
    sv->reconfigs += 1;
    sv->reconfig_s += std::chrono::duration<double>(Clock::now() - t).count();
    return 0;
}

bool same_members(soil_hip_server* sv) {
    const size_t nb = sv->batch ? sv->batch->boxes.size() : 0;
    if (nb != sv->active.size()) return false;
    for (size_t k = 0; k < nb; ++k)
        if (sv->batch->boxes[k] != sv->active[k]->box) return false;
    return true;
}

void fail(soil_hip_server* sv, int rc) {
    sv->error = rc;
    sv->error_msg = soil_hip_last_error();
}

void loop(soil_hip_server* sv) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipSetDevice(sv->ctx->device) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        fail(sv, set_error(SOIL_HIP_EDEVICE, "server: device %d set-up failed", sv->ctx->device));
        if (e0) (void)hipEventDestroy(e0);
        return;
    }
    uint32_t* d_col = nullptr;
    uint32_t* h_col = nullptr;
    size_t col_cap = 0;
    std::vector<std::unique_ptr<Job>> incoming;
    while (true) {
        int rc = retire(sv);
        if (rc) { fail(sv, rc); break; }
        for (auto& j : incoming) sv->active.push_back(std::move(j));
        incoming.clear();
        if (sv->active.empty()) {
            admit(sv, incoming);
            if (!incoming.empty()) continue;
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->pending.empty() && (sv->draining.load() || !sv->running.load())) break;
            sv->cv.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        if (!same_members(sv)) {
            rc = rebatch(sv);
            if (rc) { fail(sv, rc); break; }
            const size_t need = 4 * sv->active.size();
            if (need > col_cap) {
                if (d_col) (void)hipFree(d_col);
                if (h_col) (void)hipHostFree(h_col);
                d_col = nullptr;
                h_col = nullptr;
                col_cap = need * 2;
                if (hipMalloc(&d_col, col_cap * 4) != hipSuccess ||
                    hipHostMalloc(&h_col, col_cap * 4, hipHostMallocDefault) != hipSuccess) {
                    fail(sv, set_error(SOIL_HIP_EDEVICE, "server: state read-back buffers (%zu B)", col_cap * 4));
                    break;
                }
            }
        }
        // a chunk ends when the first active job reaches its step count
        int left = 1 << 30;
        double nparticles = 0.0;
        for (auto& j : sv->active) {
            left = std::min(left, j->steps - j->done_steps);
            nparticles += (double)j->n;
        }
        const int n = std::max(1, std::min(sv->chunk, left));
        (void)hipEventRecord(e0, sv->batch->stream);  // timing only
        rc = batch_step(sv->batch, n, sv->use_graph);
        if (rc) { fail(sv, rc); break; }
        phase_collect(sv->batch->L, d_col);
        if (hipMemcpyAsync(h_col, d_col, 4 * sv->active.size() * 4, hipMemcpyDeviceToHost, sv->batch->stream) !=
                hipSuccess ||
            hipEventRecord(e1, sv->batch->stream) != hipSuccess) {
            fail(sv, set_error(SOIL_HIP_EDEVICE, "server: state read-back after a chunk"));
            break;
        }
        // overlaps the chunk: admit queued jobs and set their boxes up
        admit(sv, incoming);
        if (hipEventSynchronize(e1) != hipSuccess) {
            fail(sv, set_error(SOIL_HIP_EDEVICE, "server: chunk did not complete"));
            break;
        }
        sv->batch->half_pending[0] = sv->batch->half_pending[1] = false;
        // a box that ran out of ghost or pair capacity inside the chunk froze
        // at that step: grow it and step it up to the others (and grow boxes
        // near their capacity before they overflow)
        rc = batch_settle(sv->batch, h_col, true);
        if (rc < 0) { fail(sv, rc); break; }
        if (rc > 0) std::memcpy(h_col, sv->batch->h_col_now, 16 * sv->active.size());
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        sv->busy_s += ms * 1e-3;
        sv->server_steps += (uint64_t)n;
        sv->chunks += 1;
        sv->active_sum += (double)sv->active.size();
        sv->particle_steps += nparticles * n;
        const double per = ms / n;
        const double tnow = sv->now();
        for (size_t k = 0; k < sv->active.size(); ++k) {
            Job* j = sv->active[k].get();
            const int took = std::min(j->steps, j->done_steps + n) - j->done_steps;
            j->done_steps += took;
            if (took > 0) {
                const double gap = 1e3 * (tnow - j->t_last);
                sv->gap_ms.push_back(gap);
                for (int q = 0; q < took; ++q) {
                    sv->step_ms.push_back(gap / took);
                    sv->gpu_step_ms.push_back(per);
                }
                j->t_last = tnow;
            }
            if (h_col[4 * k]) {  // device error bits: the job's results are invalid
                j->failed = true;
                check_err_bits(h_col[4 * k]);
            }
            if (h_col[4 * k + 3] != (uint32_t)j->done_steps) j->failed = true;  // device/host step drift
        }
    }
    for (auto& j : incoming) sv->active.push_back(std::move(j));
    if (sv->batch) {
        batch_free(sv->batch, false);
        sv->batch = nullptr;
    }
    if (d_col) (void)hipFree(d_col);
    if (h_col) (void)hipHostFree(h_col);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace

extern "C" {

int soil_hip_server_create(soil_hip_ctx* ctx, double budget_gb, int max_slots, int chunk_steps, int use_graph,
                           soil_hip_server** out) {
    if (!ctx || !out || budget_gb <= 0 || max_slots < 1 || chunk_steps < 1)
        return set_error(SOIL_HIP_EINVAL, "bad arguments");
    soil_hip_server* sv = new soil_hip_server();
    sv->ctx = ctx;
    sv->budget = (size_t)(budget_gb * 1e9);
    sv->max_slots = max_slots;
    sv->chunk = chunk_steps;
    sv->use_graph = use_graph != 0;
    sv->t0 = Clock::now();
    *out = sv;
    return 0;
}

int soil_hip_server_start(soil_hip_server* sv) {
    if (!sv) return set_error(SOIL_HIP_EINVAL, "null server");
    if (sv->running.load()) return set_error(SOIL_HIP_ESTATE, "already running");
    sv->running = true;
    sv->draining = false;
    sv->worker = std::thread(loop, sv);
    return 0;
}

int soil_hip_server_submit(soil_hip_server* sv, const soil_hip_box_desc* desc, size_t n, const double* pos,
                           const double* vel, const double* mass, const double* radius, int nsteps, double k,
                           int keep_result, int* job_id) {
    if (!sv || !desc || !pos || !mass || !radius || nsteps < 1) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    auto j = std::make_unique<Job>();
    j->desc = *desc;
    j->n = n;
    j->pos.assign(pos, pos + 3 * n);
    j->vel = vel ? std::vector<double>(vel, vel + 3 * n) : std::vector<double>(3 * n, 0.0);
    j->mass.assign(mass, mass + n);
    j->rad.assign(radius, radius + n);
    j->steps = nsteps;
    j->k = k;
    j->keep_result = keep_result != 0;
    std::lock_guard<std::mutex> lk(sv->mu);
    j->t_arrive = sv->now();
    j->id = sv->next_id++;
    if (job_id) *job_id = j->id;
    sv->pending.push_back(std::move(j));
    sv->cv.notify_one();
    return 0;
}

int soil_hip_server_submit_restart(soil_hip_server* sv, const soil_hip_box_desc* desc, const char* path, int format,
                                   int nsteps, double k, int keep_result, int* job_id) {
    if (!sv || !desc || !path || nsteps < 1) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    if (format != SOIL_HIP_RESTART_BINCODE && format != SOIL_HIP_RESTART_JSON)
        return set_error(SOIL_HIP_EINVAL, "unknown restart format %d", format);
    auto j = std::make_unique<Job>();
    j->desc = *desc;
    j->restart_in = path;
    j->restart_in_format = format;
    j->steps = nsteps;
    j->k = k;
    j->keep_result = keep_result != 0;
    std::lock_guard<std::mutex> lk(sv->mu);
    j->t_arrive = sv->now();
    j->id = sv->next_id++;
    if (job_id) *job_id = j->id;
    sv->pending.push_back(std::move(j));
    sv->cv.notify_one();
    return 0;
}

int soil_hip_server_set_output(soil_hip_server* sv, const char* dir, int format) {
    if (!sv) return set_error(SOIL_HIP_EINVAL, "null server");
    if (format != SOIL_HIP_RESTART_BINCODE && format != SOIL_HIP_RESTART_JSON)
        return set_error(SOIL_HIP_EINVAL, "unknown restart format %d", format);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->out_dir = dir ? dir : "";
    sv->out_format = format;
    return 0;
}

int soil_hip_server_evict(soil_hip_server* sv, int job_id, const char* path, int format) {
    if (!sv || !path) return set_error(SOIL_HIP_EINVAL, "null argument");
    if (format != SOIL_HIP_RESTART_BINCODE && format != SOIL_HIP_RESTART_JSON)
        return set_error(SOIL_HIP_EINVAL, "unknown restart format %d", format);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (job_id < 0 || job_id >= sv->next_id) return set_error(SOIL_HIP_EINVAL, "no job %d", job_id);
    for (auto& j : sv->done)
        if (j->id == job_id) return set_error(SOIL_HIP_ESTATE, "job %d has finished", job_id);
    for (auto it = sv->pending.begin(); it != sv->pending.end(); ++it) {
        Job* j = it->get();
        if (j->id != job_id) continue;
        if (j->box || !j->restart_in.empty()) break;  // being admitted, or its input is already a file

        // not started: its input state, as a restart at step 0
        std::vector<uint32_t> tag(j->n), type(j->n, 0u);
        for (size_t i = 0; i < j->n; ++i) tag[i] = (uint32_t)i;
        std::vector<double> force(3 * j->n, 0.0);
        soil_hip_restart r{};
        r.natoms = j->n;
        r.dt = j->desc.dt;
        r.n = j->n;
        r.tag = tag.data();
        r.type = type.data();
        r.pos = j->pos.data();
        r.vel = j->vel.data();
        r.force = force.data();
        r.mass = j->mass.data();
        r.radius = j->rad.data();
        int rc = soil_hip_restart_write(path, format, &r);
        if (rc) return rc;
        j->evicted = true;
        j->t_done = sv->now();
        sv->done.push_back(std::move(*it));
        sv->pending.erase(it);
        return 0;
    }
    // on the device (or about to be): it leaves at the next chunk boundary
    sv->evict_req[job_id] = {path, format};
    return 0;
}

// Waits until every submitted job is finished, then stops the loop.
int soil_hip_server_drain(soil_hip_server* sv) {
    if (!sv) return set_error(SOIL_HIP_EINVAL, "null server");
    sv->draining = true;
    sv->cv.notify_one();
    if (sv->worker.joinable()) sv->worker.join();
    sv->running = false;
    if (sv->error) return set_error(sv->error, "%s", sv->error_msg.c_str());
    return 0;
}

// out[0..23]: as documented with soil_hip_server_stats in soil_hip.h
int soil_hip_server_stats(soil_hip_server* sv, double* out) {
    if (!sv || !out) return set_error(SOIL_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    std::vector<double> comp, wait;
    int failed = 0;
    double tmax = 0.0, tmin = 1e300;
    int evicted = 0;
    for (auto& j : sv->done) {
        if (j->failed) {
            ++failed;
            continue;
        }
        if (j->evicted) {
            ++evicted;
            continue;
        }
        comp.push_back(j->t_done - j->t_arrive);
        wait.push_back(j->t_admit - j->t_arrive);
        tmax = std::max(tmax, j->t_done);
        tmin = std::min(tmin, j->t_arrive);
    }
    std::memset(out, 0, SOIL_HIP_SERVER_NSTATS * sizeof(double));
    out[0] = (double)comp.size();
    out[1] = failed;
    out[2] = sv->particle_steps;
    out[3] = sv->busy_s;
    out[4] = comp.empty() ? 0.0 : tmax - tmin;
    out[5] = pct(sv->step_ms, 0.50);
    out[6] = pct(sv->step_ms, 0.99);
    out[7] = pct(comp, 0.50);
    out[8] = pct(comp, 0.99);
    out[9] = pct(wait, 0.50);
    out[10] = pct(wait, 0.99);
    out[11] = sv->chunks ? sv->active_sum / (double)sv->chunks : 0.0;
    out[12] = (double)sv->reconfigs;
    out[13] = sv->reconfig_s;
    out[14] = (double)sv->budget / 1e9;
    out[15] = (double)sv->chunks;
    double sum = 0.0;
    for (double v : sv->step_ms) sum += v;
    out[16] = sv->step_ms.empty() ? 0.0 : sum / (double)sv->step_ms.size();
    out[17] = comp.empty() ? 0.0 : *std::max_element(comp.begin(), comp.end());
    out[18] = wait.empty() ? 0.0 : *std::max_element(wait.begin(), wait.end());
    out[19] = (double)sv->server_steps;
    out[20] = evicted;
    out[21] = pct(sv->gpu_step_ms, 0.50);
    out[22] = pct(sv->gpu_step_ms, 0.99);
    out[23] = pct(sv->gap_ms, 0.99);
    return 0;
}

// Final local positions/velocities (AoS-of-3) of a finished job submitted with keep_result.
int soil_hip_server_result(soil_hip_server* sv, int job_id, double* pos, double* vel, size_t cap, size_t* n) {
    if (!sv) return set_error(SOIL_HIP_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& j : sv->done) {
        if (j->id != job_id) continue;
        if (j->failed) return set_error(SOIL_HIP_EDEVICE, "job %d failed", job_id);
        if (j->evicted) return set_error(SOIL_HIP_ESTATE, "job %d was evicted", job_id);
        if (!j->keep_result) return set_error(SOIL_HIP_ESTATE, "job %d was submitted without keep_result", job_id);
        if (n) *n = j->n;
        if (cap < j->n) return set_error(SOIL_HIP_ECAPACITY, "result needs %zu atoms", j->n);
        if (pos) std::memcpy(pos, j->out_pos.data(), j->out_pos.size() * 8);
        if (vel) std::memcpy(vel, j->out_vel.data(), j->out_vel.size() * 8);
        return 0;
    }
    return set_error(SOIL_HIP_ESTATE, "job %d not finished", job_id);
}

int soil_hip_server_destroy(soil_hip_server* sv) {
    if (!sv) return 0;
    if (sv->worker.joinable()) {
        sv->draining = true;
        sv->running = false;
        sv->cv.notify_one();
        sv->worker.join();
    }
    for (auto& j : sv->active)
        if (j->box) soil_hip_box_destroy(j->box);
    for (auto& j : sv->pending)
        if (j->box) soil_hip_box_destroy(j->box);
    delete sv;
    return 0;
}

// Benchmark driver: `njobs` jobs of the served-stream workload arrive as a
// Poisson process of rate `rate` (jobs/s); a client thread submits each at its
// arrival time while the server runs. Inputs are generated before the clock
// starts. Returns the server statistics (see soil_hip_server_stats).
int soil_hip_server_bench_stream(soil_hip_ctx* ctx, int njobs, double rate, uint64_t seed, double budget_gb,
                                 int max_slots, int chunk_steps, int use_graph, double* stats_out) {  // SOIL_HIP_SERVER_NSTATS values
    if (!ctx || njobs < 1 || rate <= 0 || !stats_out) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    std::vector<double> arrival(njobs);
    std::vector<int32_t> steps(njobs);
    int rc = soil_hip_gen_arrivals(seed, njobs, rate, arrival.data(), steps.data());
    if (rc) return rc;
    struct In {
        soil_hip_box_desc d;
        size_t n;
        std::vector<double> pos, vel, mass, rad;
    };
    std::vector<In> ins(njobs);
    for (int j = 0; j < njobs; ++j) {
        In& in = ins[j];
        rc = soil_hip_gen_box(5, j, seed, &in.d, &in.n, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        in.pos.resize(3 * in.n);
        in.vel.resize(3 * in.n);
        in.mass.resize(in.n);
        in.rad.resize(in.n);
        rc = soil_hip_gen_box(5, j, seed, &in.d, &in.n, in.pos.data(), in.vel.data(), in.mass.data(), in.rad.data());
        if (rc) return rc;
    }
    soil_hip_server* sv = nullptr;
    rc = soil_hip_server_create(ctx, budget_gb, max_slots, chunk_steps, use_graph, &sv);
    if (rc) return rc;
    rc = soil_hip_server_start(sv);
    if (rc) return rc;
    std::thread client([&]() {
        const auto start = Clock::now();
        for (int j = 0; j < njobs; ++j) {
            std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(
                                                      std::chrono::duration<double>(arrival[j])));
            const In& in = ins[j];
            soil_hip_server_submit(sv, &in.d, in.n, in.pos.data(), in.vel.data(), in.mass.data(), in.rad.data(),
                                   steps[j], 1000.0, 0, nullptr);
        }
    });
    client.join();
    rc = soil_hip_server_drain(sv);
    int rc2 = soil_hip_server_stats(sv, stats_out);
    soil_hip_server_destroy(sv);
    return rc ? rc : rc2;
}

}  // extern "C"
