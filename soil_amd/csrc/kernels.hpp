// Launch-side declarations of the kernels (tile tables and phase launchers).
#pragma once
#include <hip/hip_runtime.h>
#include "internal.hpp"

namespace soil_hip {

constexpr uint32_t BLK = 256;              // threads per block (4 waves of 64)
constexpr int SCAN_PER_THREAD = 8;
constexpr uint32_t SCAN_TILE = BLK * SCAN_PER_THREAD;
constexpr uint32_t BORDER_PER = 4;                 // 64-atom chunks per wave in a border tile
constexpr uint32_t BORDER_TILE = BLK * BORDER_PER;  // atoms per border-compaction tile
constexpr uint32_t CSR1_GROUP = 64;                 // tiles per group sum of the tile-staged CSR build
constexpr uint32_t TP_GROUP = 64;                   // atom tiles per group sum of the pull transpose
constexpr int TP_MAXR = 64;                         // cell ranges a pull-transpose tile can merge

enum { SORT_SPATIAL = 0, SORT_BUILD = 1 };
enum { SCAN_CELLS_SPATIAL = 0, SCAN_CELLS_BUILD = 1, SCAN_CSR = 2, SCAN_TRANSPOSE = 3, SCAN_WAVES = 4, SCAN_REMOVE = 5 };

struct Tile {
    uint32_t box, start;
};

struct TileSet {
    void* d = nullptr;
    uint32_t n = 0;
};

// One launch context: a set of boxes (device array of descriptor pointers),
// the stream, and the tile tables of each item space.
struct Launch {
    BoxDev* const* boxes = nullptr;
    hipStream_t stream = nullptr;
    uint32_t maxgrid = 4096;
    // reverse comm: 1 = some box may have overlapping lo/hi selections in that
    // dim, so its two swaps need separate ordered launches
    uint8_t rev_split[3] = {1, 1, 1};
    int csr_mode = 3;
    // consumer transpose launches: the pull kernels for boxes with tpull set,
    // the atomic path's scan and scatter for the others (a batch may have both)
    bool tpull_any = false, tatomic_any = true;
    // some box has a non-periodic axis: its pbc may remove atoms, so it runs
    // as its own phase before the fused sort (fully periodic boxes wrap inside
    // the sort's count)
    bool any_open = false;
    TileSet t_box, t_local, t_local1, t_atom, t_atom1, t_ghost, t_cell, t_border;
    TileSet t_scan_cell, t_scan_local, t_scan_atom;
};

void phase_deform(const Launch& L);
void phase_initial(const Launch& L);
void phase_initial_decide(const Launch& L);  // initial + fused displacement check + decide
void phase_final_initial_decide(const Launch& L);  // previous step's final, then as above
void phase_decide(const Launch& L);
void phase_decide_only(const Launch& L);
void phase_kick(const Launch& L, bool with_final);  // k_initial with the displacement test, no decision
void phase_decide_fused(const Launch& L);           // k_decide with the sort bookkeeping
void phase_extent(const Launch& L);                 // locals' extents into BoxState::ext_key
void phase_sync_view(const Launch& L);              // force hooks' device views
// per-tile kinetic-energy partials into part[] (mask: one byte per local, or null)
void phase_ke(const Launch& L, const uint8_t* mask, double* part);
// *out = part[0] + ... + part[n-1] in index order, on the device
void sum_tiles(const double* part, uint32_t n, double* out, hipStream_t s);
void phase_classify(const Launch& L);               // boundary flags of the local rows into the slot scratch
void phase_zero_counts(const Launch& L);            // the spatial sort's counts for a regridded step
void phase_remove_ghosts(const Launch& L);
void phase_set_rebuild_now(const Launch& L);
void phase_zero(const Launch& L);
void phase_zero_fused(const Launch& L);  // leaves forces of consumer boxes to phase_force_fused
void phase_pbc(const Launch& L);
void phase_sort(const Launch& L);
// the fused step's pbc and spatial sort (pbc inside the sort's count; k_decide<true>
// did the pre-sort bookkeeping), then the permute and the post-sort bookkeeping
void phase_pbc_sort_fused(const Launch& L);
void phase_borders(const Launch& L);
void phase_forward(const Launch& L);
// timing events (any may be null): ev0 before the CSR count, evc after it,
// ev1 after the offset scan, ev2 after the fill
void phase_build(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2);
void phase_build_fused(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2);
// phase_borders + phase_forward + phase_build_fused, with the ghost fill,
// forward refresh and save_build in one launch; with `publish` the CSR offset
// scan also writes the build's state to the host mirror
void phase_rebuild_fused(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2, bool publish);
void phase_force(const Launch& L);
void phase_force_fused(const Launch& L);  // assigns forces of consumer boxes
void phase_reverse(const Launch& L);
void phase_final(const Launch& L);
void phase_step_end(const Launch& L);
// batches carry each step's final into the next step's kick-drift (a
// per-box flag) and apply what is left at the end of the call
void phase_initial_decide_pending(const Launch& L);
void phase_step_end_pending(const Launch& L);
void phase_final_pending(const Launch& L);
void phase_collect(const Launch& L, uint32_t* out);
void copy_probe(const double* src, double* dst, size_t n, hipStream_t s);

// Fixed workspace of the allocation-free entry points: NOALLOC_BYTES of
// device memory per GPU, from a file-scope __device__ array. Layout: one f64
// partial per 256-atom tile, the result, then one mask byte per local.
constexpr size_t NOALLOC_MAX_LOCAL = SOIL_HIP_NOALLOC_MAX_LOCAL;
constexpr size_t NOALLOC_PART_OFF = 0;
constexpr size_t NOALLOC_OUT_OFF = NOALLOC_MAX_LOCAL / BLK * sizeof(double);
constexpr size_t NOALLOC_MASK_OFF = NOALLOC_OUT_OFF + 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_MASK_OFF + NOALLOC_MAX_LOCAL;
constexpr int NOALLOC_MAX_DEVICES = 64;
// Looks up the calling thread's current device's copy of the workspace
// (hipGetSymbolAddress), once per device; 0 on success.
int scratch_init(int device);
// The workspace's device address on `device` (null before scratch_init).
void* scratch_ptr(int device);

}  // namespace soil_hip
