/*
 * soil_hip.h — C ABI of the MI355X (gfx950) particle substrate.
 *
 * Replaces, for one process and one GPU, the per-step substrate systems of
 * SOIL (SueHeir/soil 0.1.3). Each entry point names the reference function it
 * stands in for (paths relative to the reference repository root). The Rust
 * `soil_hip` plugin layer binds these symbols through an `extern "C"` block;
 * tests and the benchmark bind them through ctypes.
 *
 * Conventions
 *   - Every call returns int: 0 = ok, < 0 = error (SOIL_HIP_E*).
 *     soil_hip_last_error() returns the message of the last failure on the
 *     calling thread.
 *   - Host buffers are caller-owned and AoS-of-3 where the reference uses
 *     [f64; 3] (pos, vel, force, image). Device memory is owned by the box.
 *   - A box or batch handle is not thread-safe. Each box and each batch owns
 *     one HIP stream; calls on different handles may run on different threads.
 *   - Geometry is f64 end to end; integer outputs (bins, ghost lists, CSR) are
 *     bit-identical to the reference builder on the same inputs.
 */
#ifndef SOIL_HIP_H
#define SOIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define SOIL_HIP_OK 0
#define SOIL_HIP_EINVAL (-1)     /* bad argument */
#define SOIL_HIP_EDEVICE (-2)    /* HIP runtime / no device */
#define SOIL_HIP_ECAPACITY (-3)  /* a host buffer is too small (required size in out-param) */
#define SOIL_HIP_ESTATE (-4)     /* call out of order (e.g. step before setup) */
#define SOIL_HIP_EUNSUPPORTED (-5) /* case outside the device path (see message) */
#define SOIL_HIP_EBUDGET (-6)    /* serving: job does not fit the HBM budget */
#define SOIL_HIP_EHOOK (-7)      /* a force hook returned non-zero */

/* AtomData column flags (soil_derive/src/lib.rs:200-500,
 * docs/SOIL_ATOMDATA_CONTRACT.md:87-143): FORWARD = overwritten on ghosts by
 * forward comm, REVERSE = ghost values added onto owners by reverse comm,
 * ZERO = reset to 0 each step by zero_all_forces. */
#define SOIL_HIP_COL_FORWARD 1u
#define SOIL_HIP_COL_REVERSE 2u
#define SOIL_HIP_COL_ZERO 4u

/* CommState (soil_core/src/lib.rs:87-94) */
#define SOIL_HIP_FULL_REBUILD 0
#define SOIL_HIP_COMMUNICATE_ONLY 1

typedef struct soil_hip_ctx soil_hip_ctx;
typedef struct soil_hip_box soil_hip_box;
typedef struct soil_hip_batch soil_hip_batch;
typedef struct soil_hip_server soil_hip_server;

/* Box description: [domain] (soil_core/src/domain.rs:46-78), [neighbor]
 * (soil_core/src/neighbor.rs:44-80), Atom::dt, and the [deform] xy erate
 * (soil_deform/src/lib.rs:144-165). Field order is the C layout. */
typedef struct {
    double lo[3], hi[3];
    int32_t periodic[3];
    int32_t newton;          /* neighbor.newton (half list + reverse comm) */
    double skin_fraction;    /* neighbor.skin_fraction */
    double bin_min_size;     /* neighbor.bin_size */
    uint64_t every;          /* neighbor.every */
    int32_t check;           /* neighbor.check */
    int32_t pad0;
    uint64_t sort_every;     /* neighbor.sort_every */
    double dt;               /* Atom::dt */
    double shear_rate;       /* deform xy erate γ̇ (used when has_shear) */
    int32_t has_shear;
    int32_t pad1;
} soil_hip_box_desc;

/* ── context ─────────────────────────────────────────────────────────────── */
const char* soil_hip_last_error(void);
/* Opens `device` (ordinal). Replaces the wgpu GpuContext::new of
 * soil_gpu/src/lib.rs:53-111 for the HIP path. */
int soil_hip_ctx_create(int device, soil_hip_ctx** out);
int soil_hip_ctx_destroy(soil_hip_ctx* ctx);
/* Number of HIP devices visible (0 when none); no context needed. */
int soil_hip_device_count(int* out);

/* ── box lifecycle ───────────────────────────────────────────────────────── */
/* Creates a box: Domain (domain.rs:235-296 on one rank), Neighbor::new
 * (neighbor.rs:240-285), CommTopology (comm.rs:177-182). */
int soil_hip_box_create(soil_hip_ctx* ctx, const soil_hip_box_desc* desc, soil_hip_box** out);
int soil_hip_box_destroy(soil_hip_box* box);

/* Registers an AtomData column of `width` f64 per atom with SOIL_HIP_COL_*
 * flags; mirrors register_atom_data! + #[derive(AtomData)] (atom.rs:106-118,
 * soil_derive/src/lib.rs:200-356). Must precede soil_hip_upload. */
int soil_hip_register_column(soil_hip_box* box, int width, uint32_t flags, int* col_id);

/* Uploads n local atoms (Atom columns, atom.rs:406-431; push order = local
 * index). pos/vel/force/image are AoS-of-3; vel, force, image, tag, type may
 * be NULL (zeros; tag defaults to the index). inv_mass = 1/mass. */
int soil_hip_upload(soil_hip_box* box, size_t n, const double* pos, const double* vel, const double* force,
                    const double* mass, const double* radius, const uint32_t* tag, const uint32_t* type,
                    const int32_t* image);
/* Column data for locals [0, n): n*width f64 */
int soil_hip_upload_column(soil_hip_box* box, int col, const double* data, size_t n);

/* Device bytes of a box: *estimate = what its set-up allocates (before set-up:
 * from the uploaded atoms; what the server's HBM admission counts), and
 * *allocated = what its arenas hold now (0 before set-up; grows with the
 * ghost and pair capacities). Either pointer may be NULL. No reference
 * counterpart. */
int soil_hip_box_device_bytes(soil_hip_box* box, size_t* estimate, size_t* allocated);

/* Setup systems: neighbor_setup (neighbor.rs:604-665), comm_setup
 * (comm.rs:253-283) and setup_deform (soil_deform/src/lib.rs:346-420).
 * Sizes device buffers for the uploaded atoms. Environment variables read
 * here tune the CSR build and the consumer's transpose without changing
 * their output:
 * SOIL_HIP_CSR_MODE (0-5, default 5: one thread per row with wave-staged
 * candidates, each row's first neighbours kept in LDS, each tile's rows
 * staged densely, tile bases from per-group sums of the tiles' totals, then
 * a copy of every tile to its place; 3: the same count into per-row slots, then a scan over the rows'
 * counts and a fill; 4: each wave's rows' candidates flattened over its 64
 * lanes, accepted pairs compacted into a dense region per wave) and
 * SOIL_HIP_ROW_SLOTS (4-64, default 12: pair-arena entries kept per local
 * row: the slots of modes 2-3, a wave's region in mode 4, a tile's region in
 * mode 5; a row, wave or tile beyond it is walked again);
 * SOIL_HIP_MAX_GRID (256-65536, default 4096) caps the blocks of a box's
 * launches (on 10^6 spheres 4096 is fastest: 1024 costs 14 %, 8192 1 %);
 * SOIL_HIP_TRANSPOSE (atomic, the default: arrival slots from atomics, a
 * scan and a scatter; pull: with mode 5, each 256-atom tile gathers its
 * incoming pairs from the rows of its backward stencil without atomics;
 * auto: pull while the last build had at most 3 pairs per row, atomic above
 * 4) picks how the incoming-pair transpose is built (the consumer's and a
 * hook's that asks for it); pull and auto are opt-in until run on the
 * device. */
int soil_hip_setup(soil_hip_box* box);

/* [domain] boundary of one axis (domain.rs:27-34): 0 periodic, 1 fixed,
 * 2 shrink-wrap (the bounds follow the locals' extent plus a padding every
 * step, shrink_wrap at domain.rs:509-589). Before setup; overrides the
 * descriptor's periodic flag. Shrink-wrap boxes step alone (not in a batch). */
#define SOIL_HIP_BOUNDARY_PERIODIC 0
#define SOIL_HIP_BOUNDARY_FIXED 1
#define SOIL_HIP_BOUNDARY_SHRINK_WRAP 2
int soil_hip_set_boundary(soil_hip_box* box, int dim, int type);
/* shrink_wrap_padding; 0 = the ghost cutoff (domain.rs:518-528). */
int soil_hip_set_shrink_wrap_padding(soil_hip_box* box, double padding);

/* [deform] x / y / z (soil_deform/src/lib.rs:98-205, applied by apply_deform
 * :435-545): style 0 "erate" (value = rate, L = L0 (1 + rate dt step)),
 * 1 "vel" (value = velocity, L = L0 + velocity dt step), 2 "final" (linear
 * ramp to [target_lo, target_hi] over the stage's steps), -1 none. Before
 * setup; bounds at setup are the start. */
#define SOIL_HIP_DEFORM_NONE (-1)
#define SOIL_HIP_DEFORM_ERATE 0
#define SOIL_HIP_DEFORM_VEL 1
#define SOIL_HIP_DEFORM_FINAL 2
int soil_hip_set_axis_deform(soil_hip_box* box, int dim, int style, double value, double target_lo, double target_hi);
/* [deform] remap (default 1): affine remap of the locals when the box changes. */
int soil_hip_set_deform_remap(soil_hip_box* box, int remap);
/* Steps of the current stage (RunState cycle_count + cycle_remaining): the
 * horizon of the "final" style; default 1. */
int soil_hip_set_stage_steps(soil_hip_box* box, uint64_t nsteps);
/* Current domain bounds and sizes: lo[3], hi[3], size[3]. */
int soil_hip_get_domain(soil_hip_box* box, double* out9);

/* ── force hook (the GpuForce role, soil_gpu/src/gpu_state.rs:45-55) ────────
 * A constitutive force registered by a consumer runs in the Force phase of
 * every step, after the built-in consumer and before reverse_send_force, on
 * the box's stream, and accumulates into fx/fy/fz (zeroed at
 * zero_all_forces). With newton on, forces on ghosts reach their owners
 * through the reverse comm, so a half-list hook may scatter to j. The view
 * holds device pointers; locals are 0..nlocal, ghosts follow. `offsets` and
 * `indices` are the half CSR over the locals (row i = indices[offsets[i] ..
 * offsets[i+1])); `tstart`/`tsrc` the incoming pairs of every atom (column a
 * = tsrc[tstart[a] .. tstart[a+1]], any order), present when the hook asked
 * for SOIL_HIP_HOOK_TRANSPOSE or the consumer is on. */
#define SOIL_HIP_MAX_COLS 8
#define SOIL_HIP_HOOK_TRANSPOSE 1
typedef struct soil_hip_device_view {
    uint32_t nlocal, nghost, npairs, rebuilt; /* rebuilt: the CSR was rebuilt this step */
    uint32_t cap_local, cap_atoms, cap_pairs, ncols;
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
    const double* xr; /* {x, y, z, radius} per atom, 32 B */
    const double *mass, *inv_mass, *radius;
    const uint32_t *tag, *type;
    const int32_t* origin; /* a ghost's source atom */
    const uint32_t *offsets, *indices, *tstart, *tsrc;
    double* cols[SOIL_HIP_MAX_COLS]; /* registered AtomData columns, width f64 per atom */
    int32_t col_width[SOIL_HIP_MAX_COLS];
    uint64_t step;
} soil_hip_device_view;
/* `view` (host memory) is exact in soil_hip_step and soil_hip_run_force_hook.
 * In a batch the hook is also recorded into the captured step graph: there
 * only its cap_* fields are exact (counts are the capacities), and kernels
 * must read counts and pointers at run time from `d_view`, a device copy of
 * the view refreshed on the stream just before the hook runs (the spatial
 * sort swaps column buffers). Launch on `stream` (a hipStream_t); return 0
 * or an error code. */
typedef int (*soil_hip_force_hook)(void* user, const soil_hip_device_view* view,
                                   const soil_hip_device_view* d_view, void* stream);
/* Registers (fn != NULL) or removes the box's hook; flags: SOIL_HIP_HOOK_*.
 * Asking for the transpose must precede the first neighbor build. */
int soil_hip_set_force_hook(soil_hip_box* box, soil_hip_force_hook fn, void* user, int flags);
/* The view as it stands now (waits for the box stream). */
int soil_hip_get_device_view(soil_hip_box* box, soil_hip_device_view* out);
/* The hook alone, as a per-phase system (Force phase). */
int soil_hip_run_force_hook(soil_hip_box* box);

/* Enables the harmonic soft-sphere test consumer F = k(ri+rj-d) n̂ in the
 * Force phase of soil_hip_step (0 disables; the reference has no built-in
 * force, physics tiers provide it). */
int soil_hip_set_consumer(soil_hip_box* box, int enabled, double k);

/* ── per-phase systems (one call = one reference system on this box) ───────
 * Each enqueues on the box stream; results are read back with the getters. */
int soil_hip_apply_deform(soil_hip_box* box);          /* soil_deform/src/lib.rs:435-604 (axes, xy LE) */
int soil_hip_initial_integration(soil_hip_box* box);   /* soil_verlet/src/lib.rs:131-162 */
/* decide_rebuild (neighbor.rs:796-822, needs_rebuild :762-784,
 * displacement_exceeded :693-758). Runs only in CommunicateOnly, like its
 * run_if. *full_rebuild (may be NULL) receives the resulting state == FullRebuild. */
int soil_hip_decide_rebuild(soil_hip_box* box, int* full_rebuild);
int soil_hip_remove_ghost_atoms(soil_hip_box* box);    /* atom.rs:625-629 (FullRebuild only) */
int soil_hip_zero_all_forces(soil_hip_box* box);       /* atom.rs:631-635 */
int soil_hip_shrink_wrap(soil_hip_box* box);           /* domain.rs:509-589 (shrink-wrap axes; waits for the device) */
int soil_hip_pbc(soil_hip_box* box);                   /* domain.rs:611-729 (FullRebuild only) */
int soil_hip_sort_atoms_by_bin(soil_hip_box* box);     /* neighbor.rs:834-997 */
int soil_hip_borders(soil_hip_box* box);               /* comm.rs:682-860 (FullRebuild only) */
int soil_hip_forward_comm_borders(soil_hip_box* box);  /* comm.rs:582-592 (CommunicateOnly only) */
int soil_hip_bin_neighbor_list(soil_hip_box* box);     /* neighbor.rs:1013-1319 (FullRebuild only) */
int soil_hip_harmonic_force(soil_hip_box* box);        /* test consumer (Force phase) */
int soil_hip_reverse_send_force(soil_hip_box* box);    /* comm.rs:938-1024 (newton only) */
int soil_hip_final_integration(soil_hip_box* box);     /* soil_verlet/src/lib.rs:175-194 */
/* n full timesteps in ParticleSimScheduleSet order (schedule.rs:22-52). */
int soil_hip_step(soil_hip_box* box, int nsteps);
/* Waits for the box stream; reports any device-side error flag. */
int soil_hip_sync(soil_hip_box* box);
/* Sets the xy tilt and boundary velocity directly (a static triclinic box or a
 * test); marks bounds_changed like apply_deform (domain.rs:126-137 fields). */
int soil_hip_set_tilt(soil_hip_box* box, double xy, double boundary_vel_x);

/* ── readback (lazy; each waits for the box stream) ──────────────────────── */
/* counts[0..7] = nlocal, nghost, npairs, ncells, rebuilds, steps, nstencil_fwd, state */
int soil_hip_get_counts(soil_hip_box* box, uint64_t* counts8);
/* Neighbor::neighbor_offsets (nlocal+1) and neighbor_indices
 * (neighbor.rs:91-95). ECAPACITY when cap_idx < npairs (*n_idx = npairs). */
/* compute_ke (atom.rs:589-604): 0.5 * sum of m (vx^2 + vy^2 + vz^2) over the
 * locals (mask: nlocal bytes, non-zero = in the group; NULL = all). Each term
 * is evaluated as the reference does; the sum is a fixed-order tree (per-tile
 * on the device, tiles in order on the host), deterministic and within a few
 * ulps of the reference's sequential sum. */
int soil_hip_compute_ke(soil_hip_box* box, const uint8_t* mask, double* out);
/* soil_hip_compute_ke without any allocation (no reference counterpart: it
 * is for callers that cannot allocate, such as code under HIP graph capture
 * and tight serving loops). Same arguments and the same result, bit for bit;
 * the mask and the per-tile partial sums go to a fixed workspace of the
 * library (one per GPU, held until the process ends) instead of the box's
 * scratch or host vectors. At most SOIL_HIP_NOALLOC_MAX_LOCAL locals
 * (ECAPACITY above). Calls on one GPU share the workspace: calls that overlap
 * on the same GPU (from several threads) race. */
#define SOIL_HIP_NOALLOC_MAX_LOCAL 8388608
int soil_hip_compute_ke_noalloc(soil_hip_box* box, const uint8_t* mask, double* out);
/* classify_boundary_atoms (neighbor.rs:1321-1336): mask[i] = 1 when local row
 * i of the current list has a ghost neighbour (index >= nlocal), else 0 — the
 * boundary / interior split that lets interior forces overlap a halo exchange.
 * *n (may be NULL) receives nlocal; mask may be NULL to query it. */
int soil_hip_classify_boundary_atoms(soil_hip_box* box, uint8_t* mask, size_t cap, size_t* n);
int soil_hip_get_csr(soil_hip_box* box, uint32_t* offsets, size_t cap_off, uint32_t* indices, size_t cap_idx,
                     size_t* n_idx);
/* All atoms, locals then ghosts (cap = nlocal+nghost); any pointer may be NULL. */
int soil_hip_get_atoms(soil_hip_box* box, size_t cap, double* pos, double* vel, double* force, double* radius,
                       double* mass, double* inv_mass, uint32_t* tag, int32_t* origin, int32_t* image);
int soil_hip_get_column(soil_hip_box* box, int col, double* out, size_t cap_atoms);
/* Atom::atom_type of all atoms, locals then ghosts. */
int soil_hip_get_types(soil_hip_box* box, uint32_t* type, size_t cap_atoms);
/* Last build's bins: bin_start (ncells+1), bin_sorted_atoms, bin_atom_sorted_idx,
 * bin_atom_cell (nlocal+nghost each); neighbor.rs:149-161. */
int soil_hip_get_bins(soil_hip_box* box, uint32_t* bin_start, size_t cap_cells1, uint32_t* sorted_atoms,
                      uint32_t* sorted_idx, uint32_t* atom_cell, size_t cap_atoms);
/* Swap table (SwapData, comm.rs:114-122): 6 rows of
 * [recv_start, recv_count, nsend, to_proc, from_proc, off xyz, voff xyz]. */
int soil_hip_get_swaps(soil_hip_box* box, double* out, size_t cap_rows, size_t* nrows);
/* Grid of the current step, same layout as the oracle's 21-double grid record:
 * bin_count[3], bin_size[3], bin_origin[3], ghost_cutoff, max_cutoff,
 * bin_min_size, tilt[3], boundary_vel[3], cached_min_skin, uniform, uniform_cutoff_sq */
int soil_hip_get_grid(soil_hip_box* box, double* out21);
int soil_hip_get_stencil(soil_hip_box* box, int32_t* out, size_t cap, size_t* n);
/* Host only (no device): grid record and stencil (forward if newton, else full)
 * that neighbor_setup + compute_bin_grid(_triclinic) (neighbor.rs:311-495,
 * 604-665) give a box with these radii at tilt xy. */
int soil_hip_grid_for(const soil_hip_box_desc* desc, const double* radius, size_t n, double tilt_xy, double* out21,
                      int32_t* stencil, size_t cap, size_t* nstencil);

/* ── batched ensemble (many boxes, one launch per phase, hipGraph step) ─── */
/* Boxes must be set up and not used individually while in the batch. */
int soil_hip_batch_create(soil_hip_ctx* ctx, soil_hip_box** boxes, int nboxes, soil_hip_batch** out);
int soil_hip_batch_destroy(soil_hip_batch* batch);
/* Autotunes the batch's bit-neutral knobs: the CSR build's strategy (3: one
 * thread per row walking its candidates; 4: a wave's 64 lanes sharing its
 * rows' flattened candidates) and the blocks per launch (1024-16384), by
 * timing a few steps of each setting on the live batch between snapshots of
 * every box, which are restored after each trial (the batch state is
 * unchanged). Bin size is not a knob: the grid fixes the CSR row order. */
int soil_hip_batch_autotune(soil_hip_batch* batch);
/* What the last autotune chose and measured: out[0] blocks per launch,
 * out[1] CSR mode, out[2] ms per step of that setting, out[3] trials, then
 * per trial (mode, blocks per launch, ms per step). *n receives the length
 * (out may be NULL to query it). */
int soil_hip_batch_get_tuning(soil_hip_batch* batch, double* out, size_t cap, size_t* n);
/* n timesteps of every box; use_graph = 1 replays one captured step graph. */
int soil_hip_batch_step(soil_hip_batch* batch, int nsteps, int use_graph);
int soil_hip_batch_sync(soil_hip_batch* batch);

/* ── kernel timing (HIP events around the CSR build of a rebuild) ──────── */
/* period 0: off; n > 0: every n-th rebuild's build is timed (events cost a
 * few microseconds each, so a period above 1 keeps them out of most steps). */
int soil_hip_set_timing(soil_hip_box* box, int period);
/* out[0] = builds timed, out[1] = total ms of the CSR build (count, offset
 * scan, fill), out[2] = algorithmic bytes of those builds (every build's
 * average since timing was set, times out[0]), out[3] = total
 * ms of the fill kernel, out[4] = total ms of the count kernel */
int soil_hip_get_timing(soil_hip_box* box, double* out5);
int soil_hip_batch_set_timing(soil_hip_batch* batch, int enabled);
int soil_hip_batch_get_timing(soil_hip_batch* batch, double* out4);

/* ── served stream (continuous batching of simulation jobs) ─────────────── */
/* A server admits jobs (one box each) against an HBM budget (GB) and a slot
 * limit, keeps their state resident, and steps all admitted boxes together in
 * the batched form; membership changes at chunk boundaries of at most
 * `chunk_steps` steps. No reference counterpart: SOIL runs one App per process. */
int soil_hip_server_create(soil_hip_ctx* ctx, double budget_gb, int max_slots, int chunk_steps, int use_graph,
                           soil_hip_server** out);
int soil_hip_server_start(soil_hip_server* server);
/* Queues a job (copies the inputs); the harmonic consumer with constant k acts as its force. */
int soil_hip_server_submit(soil_hip_server* server, const soil_hip_box_desc* desc, size_t n, const double* pos,
                           const double* vel, const double* mass, const double* radius, int nsteps, double k,
                           int keep_result, int* job_id);
/* Queues a job whose atoms come from a RestartData file (read_restart
 * semantics); it runs `nsteps` more steps. */
int soil_hip_server_submit_restart(soil_hip_server* server, const soil_hip_box_desc* desc, const char* path,
                                   int format, int nsteps, double k, int keep_result, int* job_id);
/* Finished jobs write their RestartData to `dir`/restart_{step}_job{id}.bin
 * (or .json); NULL or "" stops it. */
int soil_hip_server_set_output(soil_hip_server* server, const char* dir, int format);
/* Preempts a job: a running one leaves at the next chunk boundary and its
 * state is written to `path` (total_cycle = steps taken so far); a queued one
 * is written as it was submitted. Resume it with soil_hip_server_submit_restart.
 * No reference counterpart: SOIL runs one App per process. */
int soil_hip_server_evict(soil_hip_server* server, int job_id, const char* path, int format);
/* Waits for every submitted job, then stops the loop. */
int soil_hip_server_drain(soil_hip_server* server);
/* out[SOIL_HIP_SERVER_NSTATS]: 0 jobs_done, 1 jobs_failed, 2 particle_steps,
 * 3 busy_s (device time of the chunks), 4 wall_s, 5-6 step p50/p99 (ms),
 * 7-8 completion p50/p99 (s), 9-10 queue wait p50/p99 (s), 11 mean_active_jobs,
 * 12 reconfigs, 13 reconfig_s, 14 budget_gb, 15 chunks, 16 step mean (ms),
 * 17 completion max (s), 18 wait max (s), 19 server_steps, 20 jobs_evicted,
 * 21-22 device step p50/p99 (ms), 23 update gap p99 (ms).
 * Step latency is per job and per step: the wall time between the job's
 * consecutive state updates (its admission, then each chunk that moved it),
 * re-batching, admission and read-back gaps included, over the steps that
 * update brought. The device step is a chunk's GPU time over its steps; the
 * update gap is the wall time between a job's consecutive updates. */
#define SOIL_HIP_SERVER_NSTATS 24
int soil_hip_server_stats(soil_hip_server* server, double* out);
/* Final local positions/velocities of a finished job submitted with keep_result. */
int soil_hip_server_result(soil_hip_server* server, int job_id, double* pos, double* vel, size_t cap, size_t* n);
int soil_hip_server_destroy(soil_hip_server* server);
/* Benchmark driver: njobs jobs of the served-stream workload (config 5) arrive
 * as a Poisson process of `rate` jobs/s and are submitted by a client thread. */
int soil_hip_server_bench_stream(soil_hip_ctx* ctx, int njobs, double rate, uint64_t seed, double budget_gb,
                                 int max_slots, int chunk_steps, int use_graph, double* stats_out);

/* ── restart and dump files (soil_print) ───────────────────────────────── */
/* RestartData (soil_print/src/lib.rs:317-398): bincode 1 as written by
 * write_restart_inner (:1425-1445) or serde_json. Dump frames (:1272-1380):
 * CSV text, little-endian binary or LAMMPS text of tag, type, pos, vel, force,
 * radius (+ IsGhost). */
#define SOIL_HIP_RESTART_BINCODE 0
#define SOIL_HIP_RESTART_JSON 1
#define SOIL_HIP_DUMP_TEXT 0
#define SOIL_HIP_DUMP_BINARY 1
#define SOIL_HIP_DUMP_LAMMPS 2

/* Host view of one RestartData. pos/vel/force are AoS-of-3 (n*3). buf[k] holds
 * buf_len[k] f64: AtomData extension k packed atom by atom (atom.rs:353-365). */
typedef struct {
    uint64_t natoms, total_cycle;
    double dt;
    size_t n;
    uint32_t *tag, *type;
    double *pos, *vel, *force, *mass, *radius;
    size_t nbuf;
    size_t* buf_len;
    double** buf;
} soil_hip_restart;

/* Writes `r` to `path` (no reference counterpart for the call itself: it is
 * write_restart_inner without the directory and step naming). */
int soil_hip_restart_write(const char* path, int format, const soil_hip_restart* r);
/* Reads `path`. Call with r->tag == NULL to get n, nbuf and (when buf_len is
 * given with r->nbuf slots) the buffer lengths; then again with arrays of
 * r->n atoms and r->nbuf buffers. ECAPACITY when they are too small. */
int soil_hip_restart_read(const char* path, int format, soil_hip_restart* r);
/* RestartData::from_atoms (lib.rs:352-380) of a box's locals and registered
 * columns at `step`, written to `path`. */
int soil_hip_box_write_restart(soil_hip_box* box, const char* path, int format, uint64_t step);
/* read_restart (lib.rs:1527-1620) into a new box in place of soil_hip_upload:
 * atoms, forces, AtomData columns (registered first, same widths) and dt;
 * image flags start at zero. *step (may be NULL) receives total_cycle. */
int soil_hip_box_read_restart(soil_hip_box* box, const char* path, int format, uint64_t* step);
/* dump_atoms_inner (lib.rs:1138-1210) for one rank: `path_stem` + .csv / .bin
 * / .lammpstrj; ghosts included (with IsGhost) when `ghost` is set. */
int soil_hip_box_write_dump(soil_hip_box* box, const char* path_stem, int format, uint64_t step, int ghost);

/* Streams `bytes` through a plain copy kernel `reps` times; *ms = time per
 * launch. Calibrates HBM traffic counters against a known byte count. */
int soil_hip_probe_copy(soil_hip_ctx* ctx, size_t bytes, int reps, double* ms);

/* ── synthetic inputs (host only; SplitMix64 + Box–Muller) ─────────────── */
/* Writes a box description and atoms for `config` (1..5) / `box_id`.
 * With pos == NULL only the desc and *n are produced (query size). */
int soil_hip_gen_box(int config, int box_id, uint64_t seed, soil_hip_box_desc* desc, size_t* n, double* pos,
                     double* vel, double* mass, double* radius);
/* Any box: n spheres with radii U[rmin,rmax] at packing fraction phi in a box
 * of edge ratios aspect[3], optional Gaussian blobs, N(0, vel_sigma) velocities. */
int soil_hip_gen_custom(uint64_t seed, size_t n, double rmin, double rmax, double phi, const double* aspect,
                        int nblobs, double blob_frac, double sigma, double vel_sigma, double shear_rate,
                        soil_hip_box_desc* desc, double* pos, double* vel, double* mass, double* radius);
/* Served-stream arrivals: Poisson times (s) of rate `rate` and steps ~ U{500..5000}. */
int soil_hip_gen_arrivals(uint64_t seed, int njobs, double rate, double* arrival_s, int32_t* steps);

#ifdef __cplusplus
}
#endif
#endif /* SOIL_HIP_H */
