// Internal structures of the soil_hip library (host + device).
//
// Device data layout per box (all SoA, f64 geometry):
//   atoms   [0, nlocal) locals, [nlocal, nlocal+nghost) ghosts, capacity cap_atoms
//           x y z vx vy vz fx fy fz mass inv_mass rad : f64
//           tag type : u32, origin : i32, imx imy imz : i32, gswap : u8 (ghost swap id)
//   alt_*   second buffers of the permuted columns (spatial sort gathers into
//           them, then the device-side descriptor swaps the pointers)
//   bins    bin_start[ncells+1], counts, sorted atoms, sorted idx, spos (x,y,z,r)
//   csr     offsets[nlocal+1], indices[cap_pairs]
//   consumer transpose: tstart[n+1], tsrc[P]
// The descriptor BoxDev lives in device memory and is the authority for
// pointers; kernels read it at run time, so a captured graph stays valid when
// the spatial sort swaps column buffers.
#pragma once
#include <cstddef>
#include <cstdint>

#include "soil_hip.h"

#define SOIL_MAX_STENCIL 256
#define SOIL_MAX_COLS SOIL_HIP_MAX_COLS
#define SOIL_NSWAP 6

namespace soil_hip {

// error bits raised on device
enum : uint32_t {
    ERR_PAIR_CAPACITY = 1u,
    ERR_GHOST_CAPACITY = 2u,
    ERR_ATOM_LEFT_BOX = 4u,   // (unused: atoms leaving a non-periodic face are removed)
    ERR_STENCIL = 8u,
    ERR_REMOVE_PLAN = 16u,    // a removal hole found no kept source (an internal inconsistency)
};

// Host-computed geometry of one box at one step (apply_deform + recompute_bins
// are functions of the step count only, so the host precomputes them).
struct Geom {
    double lo[3], hi[3], size[3];
    double tilt[3], bvel[3];
    double bin_size[3], inv_bin[3], bin_origin[3];
    double pbc_box[3], inv_pbc_box[3], half_pbc_box[3];
    double thr_lo[3], thr_hi[3];          // border selection thresholds (real or lamda)
    double swap_shift[SOIL_NSWAP][3];
    double swap_vshift[SOIL_NSWAP][3];
    double dt;
    double shear_rate, yc;
    double skin_fraction, sf_sq, disp_thr_sq;
    double ucut_sq;
    double ghost_cut;
    uint64_t every;
    int32_t nbin[3];
    uint32_t ncells;
    int32_t periodic[3];
    int32_t triclinic;
    int32_t pbc_all;
    int32_t check;
    int32_t uniform, newton, has_self;
    int32_t nstencil;
    int32_t init_shear, clear_last, bounds_changed, periodic_sort;
    // axis deformation this step: p = new_c + (p - old_c) * scale on the remapped axes
    int32_t remap[3], remap_any;
    double remap_old_c[3], remap_new_c[3], remap_scale[3];
    int32_t disjoint[3];
    int32_t stencil[SOIL_MAX_STENCIL];
    // stencil entries in ascending offset order; entry nstencil is the self cell (offset 0)
    int16_t sorder[SOIL_MAX_STENCIL + 1];
    // the stencil as runs of consecutive cells (offsets o, o+1, ... in stencil order):
    // their atoms are contiguous in the bin-sorted arrays
    int32_t nruns;
    int32_t run_off[SOIL_MAX_STENCIL];
    int16_t run_len[SOIL_MAX_STENCIL];
};

// Device-written per-box counters and flags (host reads them back).
struct BoxState {
    uint32_t nlocal, nghost, npairs, state;     // state: 0 FullRebuild, 1 CommunicateOnly
    uint32_t last_valid;                        // last_build_pos holds nlocal positions
    uint32_t exceed;                            // displacement flag of this step
    uint32_t do_sort, not_identity;
    uint64_t steps_since_build;
    uint64_t rebuilds, steps;
    uint32_t swap_start[SOIL_NSWAP], swap_count[SOIL_NSWAP];
    uint32_t swap_send[SOIL_NSWAP];             // selected (send_indices); a non-periodic swap receives none
    uint32_t rebuild_now;                       // this step takes the FullRebuild path
    uint32_t err;
    uint32_t ring_pos;                          // geometry slot of the current step
    uint32_t local_cells_valid;                 // locals are in cell order from this step's sort (neighbor.rs:503,540)
    uint32_t lb_fresh;                          // the fused permute already wrote this build's last_build_pos
    uint32_t final_pending;                     // batches: this step's final_integration is still to apply
    uint64_t csr_bytes, csr_builds;             // algorithmic bytes of CSR builds (timing)
    uint64_t step_limit;                        // the box stops (every kernel skips it) at this step count
    uint64_t ext_key[6];                        // locals' min x,y,z then max x,y,z as order-preserving keys
    uint64_t pub_seq;                           // publishes to the host mirror so far
    uint32_t nremove;                           // locals that left through a non-periodic face this pbc
    uint32_t removed_total;                     // locals removed so far
    uint32_t sstart_build;                      // the build's locals are in the spatial sort's cell order
                                                // with its per-cell starts (sstart) valid
    uint32_t tpull_now;                         // this build's transpose strategy (1 pull), latched by its count
};

// The fields the plain step waits on, published by kernels into pinned host
// memory that the device maps (system-scope stores, the sequence word last),
// so the host polls one word instead of copying the state back.
struct HostMirror {
    uint64_t seq;
    uint32_t state, err, nlocal, nghost, npairs, rebuild_now;
};

// Device builds see the arena pointers as global-memory pointers (address
// space 1), so kernels issue global loads, stores and atomics instead of flat
// ones; the layout is the same 64-bit pointer either way.
#if defined(__HIP_DEVICE_COMPILE__)
#define GMEM __attribute__((address_space(1)))
#else
#define GMEM
#endif

struct ColDev {
    GMEM double* data;
    GMEM double* alt;
    int32_t width;
    uint32_t flags;
};

struct BoxDev {
    GMEM double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *mass, *inv_mass, *rad;
    GMEM uint32_t *tag, *type;
    GMEM int32_t *origin, *imx, *imy, *imz;
    // spatial-sort alternates (same capacity)
    GMEM double *ax, *ay, *az, *avx, *avy, *avz, *amass, *ainv_mass, *arad;
    GMEM uint32_t *atag, *atype;
    GMEM int32_t *aimx, *aimy, *aimz;
    GMEM uint8_t* gswap;
    GMEM double* xr;           // AoS mirror (x, y, z, rad) per atom for neighbor gathers
    GMEM double *lbx, *lby, *lbz;
    // sort / bin scratch
    GMEM uint32_t* cell;       // per atom cell (locals: spatial sort result reused by the build)
    GMEM uint32_t* slot;       // atomic arrival slot within a cell
    GMEM uint32_t* tmp;        // unordered scatter
    GMEM uint32_t* sorted;     // bin_sorted_atoms (or spatial-sort perm)
    GMEM uint32_t* sidx;       // bin_atom_sorted_idx
    GMEM double* spos;         // 4 f64 per sorted slot: x y z r
    GMEM uint32_t* bin_start;  // cap_cells + 1
    GMEM uint32_t* counts;     // cap_cells + 1
    GMEM uint32_t* scan_part;  // per-tile partial sums (scan scratch)
    GMEM double* red;          // per-tile partial sums of f64 reductions (compute_ke)
    // spatial sort (sort_atoms_by_bin) keeps its own arrays, like the reference's
    // sort_cell / sort_counts / sort_perm, so the last build's bins stay readable
    GMEM uint32_t* scell;      // cap_local
    GMEM uint32_t* sstart;     // cap_cells + 1
    GMEM uint32_t* perm;       // cap_local
    // CSR
    GMEM uint32_t* offsets;    // cap_local + 1
    GMEM uint32_t* indices;    // cap_pairs
    GMEM uint32_t* ttmp;       // cap_pairs: CSR temp region (rows before placement)
    // consumer: incoming pairs per atom (transpose of the half list, any order)
    GMEM uint32_t* tcount;     // cap_atoms + 1
    GMEM uint32_t* tstart;     // cap_atoms + 1
    GMEM uint32_t* tslot;      // cap_pairs: arrival slot of each CSR entry in its column
    GMEM uint32_t* tsrc;       // cap_pairs: row atom of each incoming pair, grouped by column
    GMEM uint32_t* ttile;      // pull transpose: incoming pairs of each 256-atom tile
    GMEM uint32_t* tgroup;     // pull transpose: per group of 64 atom tiles, the sum of their totals
    // border compaction scratch: per tile counts of the two swaps
    GMEM uint32_t* ctile;      // CSR: per local tile, entries staged in the temp region (mode 0: or an
                               // overflow mark; mode 5: the tile's total)
    GMEM uint32_t* cgroup;     // CSR (mode 5): per group of 64 tiles, the sum of their totals
    GMEM uint32_t* wcount;     // CSR (mode 4): entries of each 64-row wave, scanned in place into wave bases
    GMEM uint32_t* bcnt;       // 6 per border tile: selections of the 6 swaps
    ColDev cols[SOIL_MAX_COLS];
    int32_t ncols;
    uint32_t cap_local, cap_atoms, cap_cells, cap_pairs, cap_tiles;
    GMEM BoxState* st;
    const GMEM Geom* geom_ring;  // per-step geometry ring (host-written)
    uint32_t ring_len;
    uint32_t ring_stride;   // Geom records between consecutive slots of this box
    double harmonic_k;
    double dt;            // Atom::dt (constant per box)
    int32_t consumer;
    int32_t fix_fast;     // consumer forces stay below 2^19 in magnitude (fast fixed-point rounding)
    int32_t transpose;    // build the incoming-pair transpose with the CSR (consumer or a hook that asks)
    int32_t tpull;        // transpose strategy (bit-neutral): 0 (default) arrival slots from atomics in
                          // the CSR placement, a scan and a scatter; 1 (opt-in, default build only) each
                          // 256-atom tile pulls its incoming pairs from the rows of its backward stencil
    GMEM soil_hip_device_view* view;  // device copy of the hook's view (refreshed by k_sync_view), or null
    GMEM HostMirror* mirror;          // device address of the box's host mirror, or null
    int32_t csr_mode;     // CSR strategy (bit-neutral): 0 LDS tiles, 1 two global passes, 2 one pass + row slots,
                          // 3 as 2 with wave-staged candidates, 4 flattened per-wave candidates into
                          // dense wave regions, 5 (default) mode 3's count with rows kept in LDS, each
                          // tile's rows staged densely, a scan over tile totals, a copy to their place
};

int set_error(int code, const char* fmt, ...);

}  // namespace soil_hip
