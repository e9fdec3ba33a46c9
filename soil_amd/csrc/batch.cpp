// Batched ensemble: many independent boxes step together, one launch per
// phase for all of them (the kernels walk a combined tile table), and the
// whole step is captured once into a hipGraph and replayed.
//
// No host decision sits inside a step: every phase kernel reads its box's
// device state (FullRebuild / CommunicateOnly, counts, error bits) and exits
// early when the phase does not apply, so one graph serves Lees–Edwards boxes
// (rebuild every step) and periodic boxes (trigger-driven rebuilds) alike.
// Per-step geometry comes from a ring of Geom records [K][nbox] that the host
// fills half a ring ahead of the device.

#include <algorithm>
#include <chrono>
#include <cstring>

#include "box.hpp"

using namespace soil_hip;

#define HIPCHK(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return set_error(SOIL_HIP_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace soil_hip {

static int batch_tiles(soil_hip_batch* bt) {
    TileTables tt;
    for (size_t k = 0; k < bt->boxes.size(); ++k) build_tiles(tt, (uint32_t)k, bt->boxes[k]->caps);
    uint32_t maxgrid = bt->L.maxgrid ? bt->L.maxgrid : 4096;
    const int rc = upload_tiles(tt, bt->L, &bt->arena_tiles);
    if (rc) return rc;
    bt->L.boxes = bt->d_ptrs;
    bt->L.stream = bt->stream;
    bt->L.maxgrid = maxgrid;
    bt->L.csr_mode = bt->boxes.empty() ? 5 : bt->boxes[0]->h_dev.csr_mode;
    bt->L.any_open = false;
    for (auto* b : bt->boxes)
        bt->L.any_open |= !(b->model.periodic[0] && b->model.periodic[1] && b->model.periodic[2]);
    set_transpose_launch(bt->L, bt->boxes.data(), bt->boxes.size());
    return 0;
}

// The Force phase to the step's end (consumer, hooks, reverse comm, the
// step's end with its final integration left pending) for `boxes` launched
// through L. Force hooks are called here too, so under capture their launches
// are recorded into the graph; they read counts and buffers from their device
// views, refreshed by k_sync_view just before.
static int launch_tail(const Launch& L, const std::vector<soil_hip_box*>& boxes) {
    phase_force_fused(L);
    int rc = 0;
    bool hooks = false;
    for (auto* b : boxes) hooks |= b->hook != nullptr;
    if (hooks) {
        phase_sync_view(L);
        for (auto* b : boxes) {
            if (!b->hook) continue;
            soil_hip_device_view v;
            fill_view(b, v, false, 0);
            const int r = b->hook(b->hook_user, &v, b->h_dev.view, (void*)L.stream);
            if (r && !rc) rc = set_error(SOIL_HIP_EHOOK, "force hook returned %d", r);
        }
    }
    phase_reverse(L);
    phase_step_end_pending(L);  // final_integration: the next step's kick-drift or phase_final_pending
    return rc;
}

// One step of `boxes` launched through L (the batch, or one box of it).
static int launch_step_on(const Launch& L, const std::vector<soil_hip_box*>& boxes) {
    phase_initial_decide_pending(L);  // the previous step's final, the step's deform work, the kick-drift
    bool zero = false;        // forces of boxes without a consumer, #[zero] columns
    for (auto* b : boxes) {
        zero |= !b->consumer;
        for (const auto& c : b->cols) zero |= (c.flags & SOIL_HIP_COL_ZERO) != 0;
    }
    if (zero) phase_zero_fused(L);
    phase_pbc_sort_fused(L);
    phase_rebuild_fused(L, nullptr, nullptr, nullptr, nullptr, false);
    return launch_tail(L, boxes);
}

// One step of every box.
static int launch_step(soil_hip_batch* bt) { return launch_step_on(bt->L, bt->boxes); }

static void drop_graph(soil_hip_batch* bt) {
    if (bt->graph_exec) (void)hipGraphExecDestroy(bt->graph_exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    bt->graph_exec = nullptr;
    bt->graph = nullptr;
}

static int capture_graph(soil_hip_batch* bt) {
    drop_graph(bt);
    HIPCHK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_step(bt);
    HIPCHK(hipStreamEndCapture(bt->stream, &bt->graph));
    if (rc) {
        drop_graph(bt);
        return rc;
    }
    HIPCHK(hipGraphInstantiate(&bt->graph_exec, bt->graph, nullptr, nullptr, 0));
    return 0;
}

// Grows box b's ghost capacity (`ghost`) and/or pair capacity (`pairs`, to
// at least 1.5 need_pairs), keeping its atoms; pushes the descriptor and
// rebuilds the box's own tile tables. The box's stream must be idle.
static int grow_box(soil_hip_box* b, bool ghost, bool pairs, uint64_t need_pairs) {
    if (ghost) {
        Caps nc = b->caps;
        const uint64_t ghost_cap = (uint64_t)b->caps.atoms - b->caps.local;
        nc.atoms = (uint32_t)std::min<uint64_t>((uint64_t)b->caps.local + 2 * ghost_cap + 4096, 0x7FFFFFFFull);
        nc.tiles = (nc.atoms + 63) / 64;
        int rc = alloc_atoms(b, nc, true);
        if (rc) return rc;
    }
    if (pairs) {
        const uint64_t want = std::max<uint64_t>(2ull * b->caps.pairs, need_pairs + need_pairs / 2) + 4096;
        int rc = alloc_pairs(b, (uint32_t)std::min<uint64_t>(want, 0xFFFFFFF0ull), true);
        if (rc) return rc;
    }
    int rc = push_desc(b);
    return rc ? rc : rebuild_tiles(b);
}

// A box frozen in a batch at a step whose borders or build ran out of
// capacity: grows it and redoes that step's build and tail on the box's own
// stream, as the plain step does (the phases before the build already ran).
// The freeze lowered the box's step limit on the device: `limit` (its own
// limit) is restored first, or every kernel would skip the box. Leaves the
// step finished (its final integration pending).
static int redo_failed_step(soil_hip_box* b, uint64_t limit) {
    for (int attempt = 0; attempt < 6; ++attempt) {
        int rc = pull_state(b);
        if (rc) return rc;
        if (b->h_st->step_limit != limit) {
            b->h_st->step_limit = limit;
            HIPCHK(hipMemcpy(&b->d_st->step_limit, &b->h_st->step_limit, 8, hipMemcpyHostToDevice));
        }
        const uint32_t err = b->h_st->err;
        if (err & ~(ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY)) return check_err_bits(err);
        if (err & ERR_GHOST_CAPACITY) {
            rc = grow_box(b, true, false, 0);
            if (rc) return rc;
            b->h_st->err &= ~ERR_GHOST_CAPACITY;
            HIPCHK(hipMemcpy(&b->d_st->err, &b->h_st->err, 4, hipMemcpyHostToDevice));
            phase_borders(b->L);
            phase_forward(b->L);
            phase_build_fused(b->L, nullptr, nullptr, nullptr, nullptr);
            continue;
        }
        if (err & ERR_PAIR_CAPACITY) {
            rc = grow_box(b, false, true, b->h_st->npairs);
            if (rc) return rc;
            // the build runs again from its start: it counts the rebuild and
            // flips the state to CommunicateOnly once more
            b->h_st->err &= ~ERR_PAIR_CAPACITY;
            b->h_st->rebuilds -= 1;
            b->h_st->state = 0;
            HIPCHK(hipMemcpy(b->d_st, b->h_st, sizeof(BoxState), hipMemcpyHostToDevice));
            phase_build(b->L, nullptr, nullptr, nullptr, nullptr);
            continue;
        }
        rc = launch_tail(b->L, {b});
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(b->stream));
        return 0;
    }
    return set_error(SOIL_HIP_ECAPACITY, "capacity growth of a batched box failed");
}

// Brings the batch's boxes to a consistent point after a run or a call:
// `col` holds each box's err, nghost, npairs, steps (k_collect; null: read
// them now). A box that froze on a capacity error is grown, finishes its
// failed step and is stepped alone (through the batch's geometry ring, whose
// slots for those steps are still in place) up to the steps the batch ran; a
// box past 80 % of its ghost or pair capacity is grown before it overflows.
// `call_ended`: the batch call applied its pending final integrations, so a
// box stepped here applies its own. Re-tiles the batch and drops its graph
// when a box grew. Returns 1 when it acted (bt->h_col_now then holds every
// box's counters afterwards; the stream is idle), 0 when nothing was needed.
int batch_settle(soil_hip_batch* bt, const uint32_t* col, bool call_ended) {
    const size_t nb = bt->boxes.size();
    auto need = [&](size_t k, const uint32_t* c) {
        const soil_hip_box* b = bt->boxes[k];
        const uint64_t ghost_cap = (uint64_t)b->caps.atoms - b->caps.local;
        return c[4 * k] != 0u || (uint64_t)c[4 * k + 1] * 10 > ghost_cap * 8 ||
               (uint64_t)c[4 * k + 2] * 10 > (uint64_t)b->caps.pairs * 8;
    };
    if (col) {
        bool any = false;
        for (size_t k = 0; k < nb && !any; ++k) any = need(k, col);
        if (!any) return 0;
    }
    HIPCHK(hipStreamSynchronize(bt->stream));
    bt->half_pending[0] = bt->half_pending[1] = false;
    phase_collect(bt->L, bt->d_col);  // the state after the drain
    HIPCHK(hipMemcpyAsync(bt->h_col_now, bt->d_col, 16 * nb, hipMemcpyDeviceToHost, bt->stream));
    HIPCHK(hipStreamSynchronize(bt->stream));
    const uint32_t* c = bt->h_col_now;
    bool grown = false;
    for (size_t k = 0; k < nb; ++k) {
        if (!need(k, c)) continue;
        soil_hip_box* b = bt->boxes[k];
        const uint32_t err = c[4 * k];
        if (err & ~(ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY)) return check_err_bits(err);
        grown = true;
        if (!err) {  // headroom: grow between steps
            const uint64_t ghost_cap = (uint64_t)b->caps.atoms - b->caps.local;
            int rc = grow_box(b, (uint64_t)c[4 * k + 1] * 10 > ghost_cap * 8,
                              (uint64_t)c[4 * k + 2] * 10 > (uint64_t)b->caps.pairs * 8, c[4 * k + 2]);
            if (rc) return rc;
            continue;
        }
        // frozen at a failed step: finish the step (with its step limit
        // restored), then step it alone to where the batch is
        for (int d = 0; d < 3; ++d) b->L.rev_split[d] = 1;  // separate ordered swaps: right for any geometry
        int rc = redo_failed_step(b, bt->limit[k]);
        if (!rc) rc = pull_state(b);
        if (rc) return rc;
        const uint64_t target =
            std::min<uint64_t>(bt->limit[k], (uint64_t)((int64_t)bt->steps_done + bt->step_offset[k]));
        while (b->h_st->steps < target) {
            const uint64_t before = b->h_st->steps;
            rc = launch_step_on(b->L, {b});
            if (!rc) rc = pull_state(b);
            if (!rc && b->h_st->err) {  // froze again: grow, finish that step
                rc = redo_failed_step(b, bt->limit[k]);
                if (!rc) rc = pull_state(b);
            }
            if (rc) return rc;
            if (b->h_st->steps != before + 1)
                return set_error(SOIL_HIP_ESTATE, "batched box %zu did not advance past step %llu", k,
                                 (unsigned long long)before);
        }
        if (call_ended) phase_final_pending(b->L);
        HIPCHK(hipStreamSynchronize(b->stream));
        rc = pull_state(b);
        if (rc) return rc;
        if (b->h_st->steps != target)
            return set_error(SOIL_HIP_ESTATE, "batched box %zu stepped to %llu, not %llu", k,
                             (unsigned long long)b->h_st->steps, (unsigned long long)target);
    }
    if (grown) {
        drop_graph(bt);
        int rc = batch_tiles(bt);
        if (rc) return rc;
    }
    phase_collect(bt->L, bt->d_col);
    HIPCHK(hipMemcpyAsync(bt->h_col_now, bt->d_col, 16 * nb, hipMemcpyDeviceToHost, bt->stream));
    HIPCHK(hipStreamSynchronize(bt->stream));
    return grown ? 1 : 0;
}

// Records each box's err, nghost, npairs and steps of the run just enqueued
// into the host buffer of half h (read when the host next waits on that half).
static int record_collect(soil_hip_batch* bt, uint32_t h) {
    phase_collect(bt->L, bt->d_col + 4 * bt->boxes.size() * (h + 1));
    HIPCHK(hipMemcpyAsync(bt->h_col[h], bt->d_col + 4 * bt->boxes.size() * (h + 1), 16 * bt->boxes.size(),
                          hipMemcpyDeviceToHost, bt->stream));
    return 0;
}

static int batch_run(soil_hip_batch* bt, int nsteps, bool use_graph) {
    const uint32_t nb = (uint32_t)bt->boxes.size();
    const uint32_t half = bt->K / 2;
    int done = 0;
    while (done < nsteps) {
        const uint64_t s0 = bt->steps_done;
        const uint32_t slot0 = (uint32_t)(s0 % bt->K);
        const uint32_t h = slot0 / half;
        const uint32_t run = std::min<uint32_t>((uint32_t)(nsteps - done), half - (slot0 % half));
        if (bt->half_pending[h]) {
            HIPCHK(hipEventSynchronize(bt->half_done[h]));
            bt->half_pending[h] = false;
            // a box that froze or filled up in that run is handled before its
            // geometry slots are written again
            int rc = batch_settle(bt, bt->h_col[h], false);
            if (rc < 0) return rc;
            if (use_graph && !bt->graph_exec) {
                rc = capture_graph(bt);
                if (rc) return rc;
            }
        }
        // host geometry for the run: staging rows [slot][box]
        bool regrid = false;
        for (uint32_t j = 0; j < run; ++j) {
            Geom* row = bt->h_ring + (size_t)(slot0 + j) * nb;
            for (uint32_t k = 0; k < nb; ++k) {
                HostModel& m = bt->boxes[k]->model;
                if (!m.apply_deform())
                    return set_error(SOIL_HIP_EINVAL, "box %u: [deform] an axis collapsed to zero size", k);
                m.sort_prepare();
                if (m.fill(row[k])) return set_error(SOIL_HIP_EUNSUPPORTED, "stencil exceeds %d offsets", SOIL_MAX_STENCIL);
                if (row[k].ncells > bt->boxes[k]->caps.cells) {
                    // a deformed box outgrew its cell arrays: drain the stream,
                    // grow them, then re-tile the batch and recapture its graph
                    HIPCHK(hipStreamSynchronize(bt->stream));
                    bt->half_pending[0] = bt->half_pending[1] = false;
                    int rc = ensure_cells(bt->boxes[k], row[k].ncells);
                    if (rc) return rc;
                    regrid = true;
                }
                bt->boxes[k]->last_geom = row[k];
            }
        }
        if (regrid) {
            int rt = batch_tiles(bt);
            if (rt) return rt;
            if (bt->graph_exec) {
                int rc = capture_graph(bt);
                if (rc) return rc;
            }
        }
        // reverse rounds: split a dim's swaps once any box can select an atom for both
        bool resplit = false;
        for (uint32_t j = 0; j < run; ++j) {
            const Geom* row = bt->h_ring + (size_t)(slot0 + j) * nb;
            for (uint32_t k = 0; k < nb; ++k)
                for (int d = 0; d < 3; ++d)
                    if (!row[k].disjoint[d] && !bt->L.rev_split[d]) {
                        bt->L.rev_split[d] = 1;
                        resplit = true;
                    }
        }
        if (resplit && bt->graph_exec) {
            HIPCHK(hipStreamSynchronize(bt->stream));
            int rc = capture_graph(bt);
            if (rc) return rc;
        }
        HIPCHK(hipMemcpyAsync(bt->d_ring + (size_t)slot0 * nb, bt->h_ring + (size_t)slot0 * nb,
                              sizeof(Geom) * nb * run, hipMemcpyHostToDevice, bt->stream));
        for (uint32_t j = 0; j < run; ++j) {
            if (use_graph) HIPCHK(hipGraphLaunch(bt->graph_exec, bt->stream));
            else {
                int rc = launch_step(bt);
                if (rc) return rc;
            }
        }
        if (done + (int)run >= nsteps) phase_final_pending(bt->L);  // the call's last step
        int rcc = record_collect(bt, h);
        if (rcc) return rcc;
        HIPCHK(hipEventRecord(bt->half_done[h], bt->stream));
        bt->half_pending[h] = true;
        bt->steps_done += run;
        done += run;
    }
    return 0;
}

// Groups set-up boxes into a batch: repoints every box at the batch's geometry
// ring (one small descriptor patch per box) and builds the combined tile tables.
int batch_make(soil_hip_ctx* ctx, soil_hip_box** boxes, int nboxes, soil_hip_batch** out) {
    if (!ctx || !boxes || nboxes <= 0 || !out) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    for (int k = 0; k < nboxes; ++k) {
        if (!boxes[k] || !boxes[k]->setup_done) return set_error(SOIL_HIP_ESTATE, "box %d not set up", k);
        if (boxes[k]->in_batch) return set_error(SOIL_HIP_ESTATE, "box %d already in a batch", k);
        if (boxes[k]->model.has_shrink())
            return set_error(SOIL_HIP_EUNSUPPORTED,
                             "box %d has shrink-wrap boundaries, whose bounds need the host every step: step it alone", k);
        for (int j = 0; j < k; ++j)
            if (boxes[j] == boxes[k]) return set_error(SOIL_HIP_EINVAL, "box %d listed twice", k);
        // one CSR count kernel serves the whole batch, and each box's fill
        // reads what its own strategy's count leaves behind
        if (boxes[k]->h_dev.csr_mode != boxes[0]->h_dev.csr_mode)
            return set_error(SOIL_HIP_EINVAL, "box %d was set up with another SOIL_HIP_CSR_MODE than box 0", k);
    }
    HIPCHK(hipSetDevice(ctx->device));
    soil_hip_batch* bt = new soil_hip_batch();
    bt->ctx = ctx;
    bt->boxes.assign(boxes, boxes + nboxes);
    HIPCHK(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&bt->half_done[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&bt->half_done[1], hipEventDisableTiming));
    HIPCHK(hipEventCreate(&bt->t0));
    HIPCHK(hipEventCreate(&bt->t1));
    const uint32_t nb = (uint32_t)nboxes;
    HIPCHK(hipMalloc(&bt->d_ring, sizeof(Geom) * bt->K * nb));
    HIPCHK(hipHostMalloc(&bt->h_ring, sizeof(Geom) * bt->K * nb, hipHostMallocDefault));
    std::vector<BoxDev*> ptrs(nb);
    for (uint32_t k = 0; k < nb; ++k) {
        soil_hip_box* b = boxes[k];
        HIPCHK(hipStreamSynchronize(b->stream));
        bt->saved_rings.push_back(b->h_dev.geom_ring);
        bt->saved_ring_len.push_back(b->h_dev.ring_len);
        // repoint the box at the batch ring: slot j of box k is record j*nb + k
        const Geom* ring = bt->d_ring + k;
        const uint32_t len = bt->K, stride = nb, zero = 0;
        HIPCHK(hipMemcpyAsync(&b->d_dev->geom_ring, &ring, sizeof(ring), hipMemcpyHostToDevice, bt->stream));
        HIPCHK(hipMemcpyAsync(&b->d_dev->ring_len, &len, 4, hipMemcpyHostToDevice, bt->stream));
        HIPCHK(hipMemcpyAsync(&b->d_dev->ring_stride, &stride, 4, hipMemcpyHostToDevice, bt->stream));
        HIPCHK(hipMemcpyAsync(&b->d_st->ring_pos, &zero, 4, hipMemcpyHostToDevice, bt->stream));
        HIPCHK(hipStreamSynchronize(bt->stream));
        b->h_dev.geom_ring = ring;
        b->h_dev.ring_len = len;
        b->h_dev.ring_stride = stride;
        b->in_batch = true;
        ptrs[k] = b->d_dev;
    }
    // per box: the step limit to restore after a freeze, and its step count
    // relative to the batch's (the step it must reach when the batch is at s)
    for (uint32_t k = 0; k < nb; ++k) {
        const int rc = pull_state(boxes[k]);
        if (rc) {
            batch_free(bt, false);
            return rc;
        }
        bt->limit.push_back(boxes[k]->h_st->step_limit);
        bt->step_offset.push_back((int64_t)boxes[k]->h_st->steps);
    }
    HIPCHK(hipMalloc(&bt->d_col, 3 * 16 * (size_t)nb));
    HIPCHK(hipHostMalloc(&bt->h_col[0], 16 * (size_t)nb, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&bt->h_col[1], 16 * (size_t)nb, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&bt->h_col_now, 16 * (size_t)nb, hipHostMallocDefault));
    std::memset(bt->h_col[0], 0, 16 * (size_t)nb);
    std::memset(bt->h_col[1], 0, 16 * (size_t)nb);
    HIPCHK(hipMalloc(&bt->d_ptrs, sizeof(BoxDev*) * nb));
    HIPCHK(hipMemcpy(bt->d_ptrs, ptrs.data(), sizeof(BoxDev*) * nb, hipMemcpyHostToDevice));
    bt->L.maxgrid = 4096;
    for (int d = 0; d < 3; ++d) bt->L.rev_split[d] = 0;  // raised by batch_run when a box needs it
    const int rt = batch_tiles(bt);
    if (rt) {
        batch_free(bt, true);
        return rt;
    }
    *out = bt;
    return 0;
}

// Releases a batch. With `restore`, every box goes back to its own ring and
// can be stepped alone again.
int batch_free(soil_hip_batch* bt, bool restore) {
    if (!bt) return 0;
    (void)hipStreamSynchronize(bt->stream);  // best effort: the batch goes away either way
    drop_graph(bt);
    for (size_t k = 0; k < bt->boxes.size(); ++k) {
        soil_hip_box* b = bt->boxes[k];
        if (restore && pull_desc(b) == 0 && pull_state(b) == 0) {
            b->h_dev.geom_ring = bt->saved_rings[k];
            b->h_dev.ring_len = bt->saved_ring_len[k];
            b->h_dev.ring_stride = 1;
            push_desc(b);
            b->ring_pos = 0;
            b->h_st->ring_pos = 0;
            (void)hipMemcpy(&b->d_st->ring_pos, &b->h_st->ring_pos, 4, hipMemcpyHostToDevice);
            (void)hipMemcpy(b->d_ring, &b->last_geom, sizeof(Geom), hipMemcpyHostToDevice);
            b->ring_ok.assign(b->ring_ok.size(), 0);  // slot 0 no longer holds h_ring[0]
            (void)rebuild_tiles(b);
        }
        b->in_batch = false;
    }
    if (bt->arena_tiles) (void)hipFree(bt->arena_tiles);
    if (bt->d_col) (void)hipFree(bt->d_col);
    for (auto p : {bt->h_col[0], bt->h_col[1], bt->h_col_now})
        if (p) (void)hipHostFree(p);
    if (bt->d_ptrs) (void)hipFree(bt->d_ptrs);
    if (bt->d_ring) (void)hipFree(bt->d_ring);
    if (bt->h_ring) (void)hipHostFree(bt->h_ring);
    for (auto e : bt->half_done)
        if (e) (void)hipEventDestroy(e);
    if (bt->t0) (void)hipEventDestroy(bt->t0);
    if (bt->t1) (void)hipEventDestroy(bt->t1);
    if (bt->stream) (void)hipStreamDestroy(bt->stream);
    delete bt;
    return 0;
}

int batch_step(soil_hip_batch* bt, int nsteps, bool use_graph) {
    if (!bt || nsteps < 0) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    if (use_graph && !bt->graph_exec) {
        int rc = capture_graph(bt);
        if (rc) return rc;
    }
    return batch_run(bt, nsteps, use_graph);
}

}  // namespace soil_hip

extern "C" {

int soil_hip_batch_create(soil_hip_ctx* ctx, soil_hip_box** boxes, int nboxes, soil_hip_batch** out) {
    return batch_make(ctx, boxes, nboxes, out);
}

int soil_hip_batch_destroy(soil_hip_batch* bt) {
    // a box that froze in the last call still finishes its steps before it
    // is handed back (best effort: the batch goes away either way)
    if (bt && hipSetDevice(bt->ctx->device) == hipSuccess) (void)batch_settle(bt, nullptr, true);
    return batch_free(bt, true);
}

int soil_hip_batch_step(soil_hip_batch* bt, int nsteps, int use_graph) {
    if (!bt || nsteps < 0) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(bt->ctx->device));
    return batch_step(bt, nsteps, use_graph != 0);
}

int soil_hip_batch_sync(soil_hip_batch* bt) {
    if (!bt) return set_error(SOIL_HIP_EINVAL, "null batch");
    HIPCHK(hipSetDevice(bt->ctx->device));
    const int rc = batch_settle(bt, nullptr, true);
    return rc < 0 ? rc : 0;
}

// Grows each box's cell arrays to the largest grid that its next `nsteps`
// steps take (the host model advanced on a copy), so that stepping them
// regrids nothing and keeps every arena where it is. Re-tiles the batch and
// drops its graph when a box grew.
static int pregrow_cells(soil_hip_batch* bt, int nsteps) {
    bool grown = false;
    for (auto* b : bt->boxes) {
        HostModel m = b->model;
        uint64_t maxc = 0;
        for (int s = 0; s < nsteps; ++s) {
            if (!m.apply_deform()) break;  // the real step reports the collapse
            m.sort_prepare();
            Geom g;
            if (m.fill(g)) break;
            maxc = std::max<uint64_t>(maxc, g.ncells);
        }
        if (maxc > b->caps.cells) {
            HIPCHK(hipStreamSynchronize(bt->stream));
            bt->half_pending[0] = bt->half_pending[1] = false;
            int rc = ensure_cells(b, maxc);
            if (rc) return rc;
            grown = true;
        }
    }
    if (grown) {
        drop_graph(bt);
        return batch_tiles(bt);
    }
    return 0;
}

// Bit-neutral launch-shape tuning: the grid-stride cap of every launch. Each
// candidate runs a few steps on the live batch after a device-side snapshot of
// every box (columns, lists, state, descriptor) and the host models; the
// snapshot is restored after each trial, so the batch state is unchanged.
// The cell arrays are first grown for every grid the trials reach, so no trial
// reallocates an arena (the snapshot restores arena contents, and its
// descriptor points into the arenas it was taken from); each trial checks it.
int soil_hip_batch_autotune(soil_hip_batch* bt) {
    if (!bt) return set_error(SOIL_HIP_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(bt->stream));
    bt->half_pending[0] = bt->half_pending[1] = false;
    const int trial = 4;
    {
        const int rc = pregrow_cells(bt, trial + 1);  // a warm-up step and the timed ones
        if (rc) return rc;
    }
    struct Snap {
        void* atoms = nullptr;
        void* pairs = nullptr;
        size_t atoms_bytes = 0, pairs_bytes = 0;
        void* arena_atoms = nullptr;  // the arenas the descriptor points into
        void* arena_pairs = nullptr;
        Caps caps;
        BoxDev dev;
        BoxState st;
        HostModel model;
        Geom last;
    };
    std::vector<Snap> snaps(bt->boxes.size());
    for (size_t k = 0; k < bt->boxes.size(); ++k) {
        soil_hip_box* b = bt->boxes[k];
        Snap& s = snaps[k];
        int rc = pull_desc(b);
        if (rc) return rc;
        rc = pull_state(b);
        if (rc) return rc;
        s.dev = b->h_dev;
        s.arena_atoms = b->arena_atoms;
        s.arena_pairs = b->arena_pairs;
        s.caps = b->caps;
        s.st = *b->h_st;
        s.model = b->model;
        s.last = b->last_geom;
        // the arenas' own sizes (an address-range query can report a larger
        // block the runtime suballocated the arena from, which does not
        // start at the arena)
        s.atoms_bytes = b->atoms_bytes;
        s.pairs_bytes = b->pairs_bytes;
        if (!s.atoms_bytes || !s.pairs_bytes) return set_error(SOIL_HIP_EDEVICE, "autotune: arena size unknown");
        HIPCHK(hipMalloc(&s.atoms, s.atoms_bytes));
        HIPCHK(hipMalloc(&s.pairs, s.pairs_bytes));
        HIPCHK(hipMemcpy(s.atoms, b->arena_atoms, s.atoms_bytes, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(s.pairs, b->arena_pairs, s.pairs_bytes, hipMemcpyDeviceToDevice));
    }
    const uint64_t steps_before = bt->steps_done;
    auto restore = [&]() -> int {
        HIPCHK(hipStreamSynchronize(bt->stream));
        for (size_t k = 0; k < bt->boxes.size(); ++k) {
            soil_hip_box* b = bt->boxes[k];
            Snap& s = snaps[k];
            if (b->arena_atoms != s.arena_atoms || b->arena_pairs != s.arena_pairs ||
                std::memcmp(&b->caps, &s.caps, sizeof(Caps)) != 0)
                return set_error(SOIL_HIP_ESTATE, "autotune: a trial reallocated box %zu (state not restorable)", k);
            HIPCHK(hipMemcpy(b->arena_atoms, s.atoms, s.atoms_bytes, hipMemcpyDeviceToDevice));
            HIPCHK(hipMemcpy(b->arena_pairs, s.pairs, s.pairs_bytes, hipMemcpyDeviceToDevice));
            b->h_dev = s.dev;
            int rc = push_desc(b);
            if (rc) return rc;
            *b->h_st = s.st;
            HIPCHK(hipMemcpy(b->d_st, &s.st, sizeof(BoxState), hipMemcpyHostToDevice));
            b->model = s.model;
            b->last_geom = s.last;
        }
        bt->steps_done = steps_before;
        bt->half_pending[0] = bt->half_pending[1] = false;
        return 0;
    };
    // the CSR strategy of every box (descriptor field and snapshot alike) and the launch context
    auto set_mode = [&](int mode) -> int {
        for (size_t k = 0; k < bt->boxes.size(); ++k) {
            soil_hip_box* b = bt->boxes[k];
            b->h_dev.csr_mode = mode;
            snaps[k].dev.csr_mode = mode;
            HIPCHK(hipMemcpy(&b->d_dev->csr_mode, &mode, sizeof(int), hipMemcpyHostToDevice));
            b->L.csr_mode = mode;
            set_transpose_launch(b->L, &b, 1);  // the pull transpose follows the default build only
        }
        bt->L.csr_mode = mode;
        set_transpose_launch(bt->L, bt->boxes.data(), bt->boxes.size());
        return 0;
    };
    auto trial_ms = [&](uint32_t mg, int mode, double& ms) -> int {
        int rc = set_mode(mode);
        if (rc) return rc;
        bt->L.maxgrid = mg;
        drop_graph(bt);
        rc = capture_graph(bt);
        if (rc) return rc;
        rc = batch_run(bt, 1, true);  // warm
        if (rc < 0) return rc;
        HIPCHK(hipStreamSynchronize(bt->stream));
        auto t0 = std::chrono::steady_clock::now();
        rc = batch_run(bt, trial, true);
        if (rc < 0) return rc;
        HIPCHK(hipStreamSynchronize(bt->stream));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial;
        bt->tune_trials.push_back({(double)mode, (double)mg, ms});
        return restore();
    };
    auto free_snaps = [&]() {
        for (auto& s : snaps) {
            (void)hipFree(s.atoms);
            (void)hipFree(s.pairs);
        }
    };
    bt->tune_trials.clear();
    // coordinate descent: the CSR strategy at the current launch cap (mode 5:
    // rows staged per 256-row tile, then placed at the tiles' bases; mode 3:
    // one thread per row into row slots, then an offset scan and a fill;
    // mode 4: a wave's 64 lanes sharing its rows' flattened candidates), then
    // the launch cap. The strategies are timed in two interleaved rounds and
    // each keeps its faster round, so that no strategy's figure rests on one
    // few-step sample taken first or last.
    const int modes[3] = {5, 4, 3};
    const uint32_t cands[5] = {1024, 2048, 4096, 8192, 16384};
    const uint32_t mg0 = bt->L.maxgrid;
    int best_mode = bt->L.csr_mode;
    uint32_t best = mg0;
    double best_ms = 1e30;
    double mode_ms[3] = {1e30, 1e30, 1e30};
    for (int round = 0; round < 2; ++round) {
        for (int k = 0; k < 3; ++k) {
            double ms = 0.0;
            const int rc = trial_ms(mg0, modes[k], ms);
            if (rc) {
                free_snaps();
                return rc;
            }
            mode_ms[k] = std::min(mode_ms[k], ms);
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (mode_ms[k] < best_ms) {
            best_ms = mode_ms[k];
            best_mode = modes[k];
        }
    }
    for (uint32_t mg : cands) {
        if (mg == mg0) continue;
        double ms = 0.0;
        const int rc = trial_ms(mg, best_mode, ms);
        if (rc) {
            free_snaps();
            return rc;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = mg;
        }
    }
    free_snaps();
    int rc = set_mode(best_mode);
    if (rc) return rc;
    bt->L.maxgrid = best;
    bt->tune_ms = best_ms;
    drop_graph(bt);
    return 0;
}

int soil_hip_batch_get_tuning(soil_hip_batch* bt, double* out, size_t cap, size_t* n) {
    if (!bt) return set_error(SOIL_HIP_EINVAL, "null batch");
    const size_t need = 4 + 3 * bt->tune_trials.size();
    if (n) *n = need;
    if (!out) return 0;
    if (cap < need) return set_error(SOIL_HIP_ECAPACITY, "tuning record needs %zu values", need);
    out[0] = (double)bt->L.maxgrid;
    out[1] = (double)bt->L.csr_mode;
    out[2] = bt->tune_ms;
    out[3] = (double)bt->tune_trials.size();
    for (size_t t = 0; t < bt->tune_trials.size(); ++t)
        for (int q = 0; q < 3; ++q) out[4 + 3 * t + q] = bt->tune_trials[t][q];
    return 0;
}

int soil_hip_batch_set_timing(soil_hip_batch* bt, int enabled) {
    if (!bt) return set_error(SOIL_HIP_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(bt->stream));
    bt->timing = enabled != 0;
    bt->t_ms = 0.0;
    uint64_t bytes = 0;
    for (auto* b : bt->boxes) {
        int rc = pull_state(b);
        if (rc) return rc;
        bytes += b->h_st->csr_bytes;
    }
    bt->bytes_at_enable = bytes;
    return 0;
}

int soil_hip_batch_get_timing(soil_hip_batch* bt, double* out4) {
    if (!bt || !out4) return set_error(SOIL_HIP_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(bt->stream));
    uint64_t bytes = 0, builds = 0;
    for (auto* b : bt->boxes) {
        int rc = pull_state(b);
        if (rc) return rc;
        bytes += b->h_st->csr_bytes;
        builds += b->h_st->csr_builds;
    }
    out4[0] = (double)builds;
    out4[1] = bt->t_ms;
    out4[2] = (double)(bytes - bt->bytes_at_enable);
    out4[3] = 0.0;
    return 0;
}

}  // extern "C"
