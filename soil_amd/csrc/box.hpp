// Host-side box / context / batch objects.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "kernels.hpp"
#include "soil_hip.h"

struct soil_hip_ctx {
    int device = 0;
};

namespace soil_hip {

struct HostCol {
    int width;
    uint32_t flags;
    std::vector<double> data;  // locals, uploaded at setup
};

struct TileTables {
    std::vector<Tile> box, local, local1, atom, atom1, ghost, cell, border, scan_cell, scan_local, scan_atom;
};

struct Caps {
    uint32_t local = 0, atoms = 0, cells = 0, pairs = 0, tiles = 0;
};

void build_tiles(TileTables& tt, uint32_t box_index, const Caps& c);
int upload_tiles(const TileTables& tt, Launch& L, void** arena);

}  // namespace soil_hip

struct soil_hip_box;
struct soil_hip_batch;
namespace soil_hip {
int alloc_atoms(soil_hip_box* b, const Caps& nc, bool keep);
int alloc_pairs(soil_hip_box* b, uint32_t cap_pairs, bool keep);
uint32_t row_slots_setting();
void set_transpose_launch(Launch& L, soil_hip_box* const* boxes, size_t n);
int tune_transpose(soil_hip_box* b);
uint32_t pair_capacity(uint64_t cap_pairs, uint32_t row_slots, uint32_t cap_local);
uint32_t initial_pair_capacity(uint32_t row_slots, uint32_t cap_local);
// Device bytes a box will hold after set-up (its arenas at their initial capacities).
size_t box_bytes_estimate(soil_hip_box* b);
// Device bytes soil_hip_setup allocated for the box's arenas (0 before set-up).
size_t box_bytes_allocated(soil_hip_box* b);
int push_desc(soil_hip_box* b);
int pull_desc(soil_hip_box* b);
int pull_state(soil_hip_box* b);
int rebuild_tiles(soil_hip_box* b);
int ensure_cells(soil_hip_box* b, uint64_t ncells);
// device view for a force hook; exact = host counts are current (plain step)
void fill_view(soil_hip_box* b, soil_hip_device_view& v, bool exact, uint64_t step);
int call_hook(soil_hip_box* b, bool exact, uint64_t step);
int check_err_bits(uint32_t err);
int batch_make(soil_hip_ctx* ctx, soil_hip_box** boxes, int nboxes, soil_hip_batch** out);
int batch_free(soil_hip_batch* bt, bool restore);
int batch_step(soil_hip_batch* bt, int nsteps, bool use_graph);
// After a batch call or a run: grows boxes near their capacity, and grows,
// finishes and steps up to the batch every box that froze on a capacity
// error (col: err, nghost, npairs, steps per box, or null to read them).
// 1 when it acted (bt->h_col_now then holds the counters), 0 when not, < 0 on error.
int batch_settle(soil_hip_batch* bt, const uint32_t* col, bool call_ended);
}  // namespace soil_hip

struct soil_hip_box {
    soil_hip_ctx* ctx = nullptr;
    soil_hip_box_desc desc{};
    soil_hip::HostModel model;
    std::vector<soil_hip::HostCol> cols;
    // host copies of the uploaded locals (SoA), consumed by setup
    std::vector<double> hx, hy, hz, hvx, hvy, hvz, hfx, hfy, hfz, hmass, hrad;
    std::vector<uint32_t> htag, htype;
    std::vector<int32_t> himx, himy, himz;
    size_t nlocal = 0;
    bool setup_done = false;
    bool in_batch = false;
    // device
    hipStream_t stream = nullptr;
    soil_hip::Caps caps;
    void* arena_atoms = nullptr;
    void* arena_pairs = nullptr;
    void* arena_tiles = nullptr;
    size_t atoms_bytes = 0, pairs_bytes = 0, tiles_bytes = 0;  // sizes of the three arenas
    size_t est_bytes = 0;  // box_bytes_estimate at set-up
    // entries per local row in the pair arena's row slots: the one-pass CSR
    // count stores a row's first row_slots pairs there and the fill walks a
    // longer row again on one lane, so the width sets how rarely a wave waits
    // on such a lane (1M spheres at phi 0.5: 10^4 rows exceed 8, 254 exceed 12)
    uint32_t row_slots = 12;
    // consumer transpose strategy (SOIL_HIP_TRANSPOSE: atomic (default), pull,
    // auto): auto pulls while the last build had at most 3 pairs per local row
    // and keeps the atomics above 4; tpull_ok: the box's stencil has few
    // enough runs for the pull's merged ranges
    int tmode = 0;
    bool tpull_ok = false;
    soil_hip::BoxDev h_dev{};        // host mirror of the device descriptor
    soil_hip::BoxDev* d_dev = nullptr;
    soil_hip::BoxDev** d_ptrs = nullptr;  // one-entry pointer array for single-box launches
    soil_hip::BoxState* d_st = nullptr;
    soil_hip::BoxState* h_st = nullptr;   // pinned
    soil_hip_device_view* d_view = nullptr;  // device copy of the force hook's view
    soil_hip::HostMirror* h_mirror = nullptr;  // pinned, device-mapped (fields published by kernels)
    uint64_t pub_expect = 0;                   // publishes enqueued so far
    uint32_t rebuild_hist = 0;                 // plain stepping: bit 0 the last step rebuilt, bit 1 the one before
    std::vector<uint8_t> ring_ok;              // device ring slot k holds h_ring[k]
    soil_hip::Geom* d_ring = nullptr;
    soil_hip::Geom* h_ring = nullptr;     // pinned
    uint32_t ring_len = 64;
    uint32_t ring_pos = 0;                // host mirror of st.ring_pos
    soil_hip::Geom last_geom{};
    soil_hip::Launch L;
    bool consumer = false;
    double harmonic_k = 0.0;
    double max_radius = 0.0;  // of the uploaded atoms (bounds the consumer's forces)
    // force hook (Force phase, after the consumer)
    soil_hip_force_hook hook = nullptr;
    void* hook_user = nullptr;
    int hook_flags = 0;
    // timing
    uint32_t timing = 0;       // time every timing-th rebuild's CSR build (0: off)
    uint64_t timing_ctr = 0;   // rebuilds launched since timing was set
    uint64_t builds_at_enable = 0;
    std::vector<hipEvent_t> ev_pool;
    struct Pending { hipEvent_t e0, ec, e1, e2; };
    std::vector<Pending> pending;
    double t_builds = 0.0, t_csr_ms = 0.0, t_fill_ms = 0.0, t_count_ms = 0.0;
    uint64_t bytes_at_enable = 0;
};

struct soil_hip_batch {
    soil_hip_ctx* ctx = nullptr;
    std::vector<soil_hip_box*> boxes;
    hipStream_t stream = nullptr;
    soil_hip::BoxDev** d_ptrs = nullptr;
    void* arena_tiles = nullptr;
    size_t atoms_bytes = 0, pairs_bytes = 0, tiles_bytes = 0;  // sizes of the three arenas
    size_t est_bytes = 0;  // box_bytes_estimate at set-up
    soil_hip::Launch L;
    soil_hip::Geom* d_ring = nullptr;  // [nbox][K]
    soil_hip::Geom* h_ring = nullptr;  // pinned [2][nbox][K/2] staging halves
    uint32_t K = 64;
    uint64_t steps_done = 0;
    hipEvent_t half_done[2] = {nullptr, nullptr};
    bool half_pending[2] = {false, false};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<const soil_hip::Geom*> saved_rings;
    std::vector<uint32_t> saved_ring_len;
    // per box: err, nghost, npairs, steps after a run (k_collect): [0] now, [1..2] halves
    uint32_t* d_col = nullptr;
    uint32_t* h_col[2] = {nullptr, nullptr};
    uint32_t* h_col_now = nullptr;
    std::vector<uint64_t> limit;        // each box's step limit (restored after a freeze)
    std::vector<int64_t> step_offset;   // a box's steps when the batch has run none
    std::vector<std::array<double, 3>> tune_trials;  // autotune: (csr mode, blocks per launch, ms per step)
    double tune_ms = 0.0;                             // the chosen setting's ms per step (0: not tuned)
    bool timing = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double t_ms = 0.0;
    uint64_t bytes_at_enable = 0;
};
