// HIP kernels for gfx950 (wave64). Every kernel walks a tile table
// (box, first item) with a grid-stride loop, so one launch covers one box (the
// plain path) or a whole ensemble (the batched path) with the same code.
//
// Arithmetic follows the reference expression for expression (f64, no
// contraction: the library is built with -ffp-contract=off; the only fused
// multiply-adds are explicit fma() where the reference writes mul_add).
// Integer outputs are order-exact: every sort is a stable counting sort
// (atomic arrival slots, then an in-segment rank by index), every ghost append
// is an ordered wave-ballot compaction.

#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace soil_hip {

// Tile iteration. Blocks are dispatched round-robin over the 8 XCDs. Each
// window of gridDim.x consecutive tiles is cut into 8 contiguous chunks, one
// per XCD (block b takes chunk b % 8), so the neighbor cells of a tile sit in
// the same XCD's L2 while every XCD still gets work when only a prefix of a
// capacity-sized tile table is active.
struct TileIter {
    uint32_t t, hi, step;
    // spread: plain block order (see FOR_TILES_SPREAD)
    __device__ __forceinline__ explicit TileIter(uint32_t n, bool spread = false) {
        const uint32_t g = gridDim.x;
        hi = n;
        step = g;
        if ((g & 7u) == 0u && !spread) t = (blockIdx.x & 7u) * (g >> 3) + (blockIdx.x >> 3);
        else t = blockIdx.x;
    }
};
#define FOR_TILES(t) \
    for (TileIter it_(ntiles); it_.t < it_.hi; it_.t += it_.step) \
        if (const uint32_t t = it_.t; true)
// Tiles in block order, which deals consecutive tiles to different XCDs: for
// work that only the last tiles of a box have (the ghosts), which the XCD
// windows of FOR_TILES would leave to one XCD.
#define FOR_TILES_SPREAD_IF(t, spread) \
    for (TileIter it_(ntiles, spread); it_.t < it_.hi; it_.t += it_.step) \
        if (const uint32_t t = it_.t; true)
#define FOR_TILES_SPREAD(t) FOR_TILES_SPREAD_IF(t, true)
#define KARGS BoxDev* const* __restrict__ boxes, const Tile* __restrict__ tiles, uint32_t ntiles

// Kernels read box descriptors and the step geometry through the constant
// address space: both are fixed while a kernel runs (the host writes the
// geometry between launches; only k_swap_ptrs and k_post_sort rewrite a
// descriptor, through BoxDev&). Stores to the arenas cannot alias them, so
// their fields load once, into scalar registers, instead of being re-read
// after every store.
#if defined(__HIP_DEVICE_COMPILE__)
#define CMEM __attribute__((address_space(4)))
#else
#define CMEM
#endif
typedef const CMEM BoxDev CBox;
typedef const CMEM Geom CGeom;
__device__ __forceinline__ CBox& box_at(BoxDev* const* boxes, uint32_t k) { return *(CBox*)boxes[k]; }

// A box whose step count reached its limit is frozen: every phase skips it.
__device__ __forceinline__ bool live(CBox& b) { return b.st->steps < b.st->step_limit; }

__device__ __forceinline__ CGeom& geom_of(CBox& b) {
    return *(CGeom*)(b.geom_ring + (size_t)b.st->ring_pos * b.ring_stride);
}
#if defined(__HIP_DEVICE_COMPILE__)
// the two kernels that rewrite a descriptor hold it as BoxDev&
__device__ __forceinline__ bool live(const BoxDev& b) { return b.st->steps < b.st->step_limit; }
__device__ __forceinline__ CGeom& geom_of(const BoxDev& b) {
    return *(CGeom*)(b.geom_ring + (size_t)b.st->ring_pos * b.ring_stride);
}
#endif

// neighbor.rs:915-922 / 1065-1076: floor, saturating `as i32`, clamp.
__device__ __forceinline__ int cell_coord(double v, int n) {
    double f = floor(v);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)(n - 1)) return n - 1;
    return (int)f;
}

// domain.rs:211-217
__device__ __forceinline__ void x2lamda(CGeom& g, double x, double y, double z, double& lx, double& ly,
                                        double& lz) {
    double xy = g.tilt[0], xz = g.tilt[1], yz = g.tilt[2];
    lz = (z - g.lo[2]) / g.size[2];
    ly = (y - g.lo[1] - yz * lz) / g.size[1];
    lx = (x - g.lo[0] - xy * ly - xz * lz) / g.size[0];
}

// domain.rs:222-229
__device__ __forceinline__ void lamda2x(CGeom& g, double lx, double ly, double lz, double& x, double& y,
                                        double& z) {
    double xy = g.tilt[0], xz = g.tilt[1], yz = g.tilt[2];
    x = g.lo[0] + g.size[0] * lx + xy * ly + xz * lz;
    y = g.lo[1] + g.size[1] * ly + yz * lz;
    z = g.lo[2] + g.size[2] * lz;
}

__device__ __forceinline__ uint32_t cell_of(CGeom& g, double x, double y, double z) {
    double px = x, py = y, pz = z;
    if (g.triclinic) x2lamda(g, x, y, z, px, py, pz);
    int cx = cell_coord((px - g.bin_origin[0]) * g.inv_bin[0], g.nbin[0]);
    int cy = cell_coord((py - g.bin_origin[1]) * g.inv_bin[1], g.nbin[1]);
    int cz = cell_coord((pz - g.bin_origin[2]) * g.inv_bin[2], g.nbin[2]);
    return (uint32_t)(cx * g.nbin[1] * g.nbin[2] + cy * g.nbin[2] + cz);
}

// Sets *flag to 1 if any active lane of the wave has pred; one atomic per wave
// at most, and none once the flag is already set.
__device__ __forceinline__ void wave_flag(uint32_t* flag, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0ull) return;
    const uint64_t act = __ballot(1);
    if ((uint32_t)__lane_id() == (uint32_t)__builtin_ctzll(act)) {
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) atomicOr(flag, 1u);
    }
}

__device__ __forceinline__ void put_xr(CBox& b, uint32_t i, double x, double y, double z) {
    double* p = b.xr + 4 * (size_t)i;
    p[0] = x;
    p[1] = y;
    p[2] = z;
}

__device__ __forceinline__ uint32_t lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Block-wide sum of one u64 per thread (BLK threads, wave64), in every thread.
__device__ uint64_t block_sum64(uint64_t v, uint64_t* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0u) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < BLK / 64; ++w) tot += lds[w];
    __syncthreads();  // lds is reused by the caller's next call
    return tot;
}

// Block-wide exclusive scan of one u32 per thread (BLK threads, wave64).
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) lds[wid] = x;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < BLK / 64; ++w) {
        uint32_t s = lds[w];
        if (w < wid) wpre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wpre + x - v;
}

// ── integration: soil_verlet/src/lib.rs:131-194 ───────────────────────────
__global__ void __launch_bounds__(BLK) k_final(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= b.st->nlocal) continue;
        // the box's dt, not the ring's: a final carried past the step's end must
        // not read the next step's geometry slot
        const double h = 0.5 * b.dt * b.inv_mass[i];
        b.vx[i] = b.vx[i] + h * b.fx[i];
        b.vy[i] = b.vy[i] + h * b.fy[i];
        b.vz[i] = b.vz[i] + h * b.fz[i];
    }
}
// The pending finals at the end of a batch call, also of boxes that reached
// their step limit during it (k_clear_pending then clears the flags)
__global__ void __launch_bounds__(BLK) k_final_pending(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!b.st->final_pending) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= b.st->nlocal) continue;
        const double h = 0.5 * b.dt * b.inv_mass[i];
        b.vx[i] = b.vx[i] + h * b.fx[i];
        b.vy[i] = b.vy[i] + h * b.fy[i];
        b.vz[i] = b.vz[i] + h * b.fz[i];
    }
}
__global__ void k_clear_pending(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x == 0) box_at(boxes, tiles[t].box).st->final_pending = 0u;
    }
}

// apply_deform's per-atom work (soil_deform/src/lib.rs:522-534, 568-574): the
// affine remap of the deformed axes, p = new_c + (p - old_c) * scale, then the
// one-time streaming profile v_x += γ̇ (y - y_c) on the remapped positions.
__global__ void __launch_bounds__(BLK) k_deform_init(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        CGeom& g = geom_of(b);
        if (!g.init_shear && !g.remap_any) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= b.st->nlocal) continue;
        if (g.remap_any) {
            double p[3] = {b.x[i], b.y[i], b.z[i]};
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (g.remap[k]) p[k] = g.remap_new_c[k] + (p[k] - g.remap_old_c[k]) * g.remap_scale[k];
            b.x[i] = p[0];
            b.y[i] = p[1];
            b.z[i] = p[2];
            put_xr(b, i, p[0], p[1], p[2]);
        }
        if (g.init_shear) b.vx[i] += g.shear_rate * (b.y[i] - g.yc);
    }
}

// displacement_exceeded (neighbor.rs:693-758) applies in CommunicateOnly with
// a valid last build, unless `every` is set without `check`
__device__ __forceinline__ bool disp_active(CGeom& g, const BoxState* st) {
    return st->state == 1u && st->last_valid && !g.clear_last && (g.every == 0 || g.check);
}

__device__ __forceinline__ bool disp_exceeds(CGeom& g, CBox& b, uint32_t i, double x, double y,
                                             double z) {
    double dx = x - b.lbx[i];
    double dy = y - b.lby[i];
    double dz = z - b.lbz[i];
    if (g.pbc_all) {
        dx -= g.pbc_box[0] * round(dx * g.inv_pbc_box[0]);
        dy -= g.pbc_box[1] * round(dy * g.inv_pbc_box[1]);
        dz -= g.pbc_box[2] * round(dz * g.inv_pbc_box[2]);
        return fma(dx, dx, fma(dy, dy, dz * dz)) > g.disp_thr_sq;
    }
    if (g.periodic[0]) {
        if (dx > g.half_pbc_box[0]) dx -= g.pbc_box[0];
        else if (dx < -g.half_pbc_box[0]) dx += g.pbc_box[0];
    }
    if (g.periodic[1]) {
        if (dy > g.half_pbc_box[1]) dy -= g.pbc_box[1];
        else if (dy < -g.half_pbc_box[1]) dy += g.pbc_box[1];
    }
    if (g.periodic[2]) {
        if (dz > g.half_pbc_box[2]) dz -= g.pbc_box[2];
        else if (dz < -g.half_pbc_box[2]) dz += g.pbc_box[2];
    }
    return dx * dx + dy * dy + dz * dz > g.disp_thr_sq;
}

// initial_integration (soil_verlet/src/lib.rs:131-162). With DISP (the fused
// step) the step's apply_deform work comes first (k_deform_init's remap and
// one-time shear profile, in registers), and the displacement check of the
// following decide_rebuild runs on the new positions while they are in
// registers (nothing between the two systems changes positions or the state
// it depends on).
// With FINAL the previous step's final_integration (v += h f) is applied
// first of all, in registers: within a multi-step call nothing between the
// two steps reads or writes velocities or forces, and the order final ->
// deform -> initial is the reference's. FINAL: 0 none, 1 always, 2 when the
// box's final is pending (batches; k_decide clears the flag).
template <bool DISP, int FINAL>
__global__ void __launch_bounds__(BLK) k_initial(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        CGeom& g = geom_of(b);
        const uint32_t i = tl.start + threadIdx.x;
        const bool act = i < st->nlocal;
        bool ex = false;
        if (act) {
            const double dt = g.dt;
            const double h = 0.5 * dt * b.inv_mass[i];
            const double fx = b.fx[i], fy = b.fy[i], fz = b.fz[i];
            double u0 = b.vx[i], u1 = b.vy[i], u2 = b.vz[i];
            double p0 = b.x[i], p1 = b.y[i], p2 = b.z[i];
            if (FINAL == 1 || (FINAL == 2 && st->final_pending != 0u)) {
                u0 = u0 + h * fx;
                u1 = u1 + h * fy;
                u2 = u2 + h * fz;
            }
            if (DISP && g.remap_any) {  // as k_deform_init
                if (g.remap[0]) p0 = g.remap_new_c[0] + (p0 - g.remap_old_c[0]) * g.remap_scale[0];
                if (g.remap[1]) p1 = g.remap_new_c[1] + (p1 - g.remap_old_c[1]) * g.remap_scale[1];
                if (g.remap[2]) p2 = g.remap_new_c[2] + (p2 - g.remap_old_c[2]) * g.remap_scale[2];
            }
            if (DISP && g.init_shear) u0 += g.shear_rate * (p1 - g.yc);
            const double v0 = u0 + h * fx;
            const double v1 = u1 + h * fy;
            const double v2 = u2 + h * fz;
            b.vx[i] = v0;
            b.vy[i] = v1;
            b.vz[i] = v2;
            const double x = p0 + v0 * dt;
            const double y = p1 + v1 * dt;
            const double z = p2 + v2 * dt;
            b.x[i] = x;
            b.y[i] = y;
            b.z[i] = z;
            put_xr(b, i, x, y, z);
            if (DISP && disp_active(g, st)) ex = disp_exceeds(g, b, i, x, y, z);
        }
        if (DISP) {
            // the fused step's spatial sort counts from zero (k_sort_zero's work,
            // done here for every step; the previous step's build no longer
            // needs the counts)
            const uint32_t stride = ((b.cap_local + BLK - 1u) / BLK) * BLK;
            for (uint32_t c = i; c <= g.ncells; c += stride) b.counts[c] = 0u;
            wave_flag(&st->exceed, ex);
        }
    }
}
__global__ void __launch_bounds__(BLK) k_disp(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        CGeom& g = geom_of(b);
        if (!disp_active(g, st)) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= st->nlocal) continue;
        wave_flag(&st->exceed, disp_exceeds(g, b, i, b.x[i], b.y[i], b.z[i]));
    }
}
// One thread of the box: the state fields the plain step's host waits on,
// into the host mirror; the sequence word goes last, at system scope.
// err and npairs come from the caller when it has just changed them itself.
__device__ __forceinline__ void publish_vals(CBox& b, uint32_t err, uint32_t npairs) {
    GMEM HostMirror* m = b.mirror;
    if (!m) return;
    GMEM BoxState* st = b.st;
    const uint64_t seq = st->pub_seq + 1u;
    st->pub_seq = seq;
    m->state = st->state;
    m->err = err;
    m->nlocal = st->nlocal;
    m->nghost = st->nghost;
    m->npairs = npairs;
    m->rebuild_now = st->rebuild_now;
    __hip_atomic_store(&m->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void publish(CBox& b) { publish_vals(b, b.st->err, b.st->npairs); }

// exceed: some local moved past the threshold since the last build
template <bool REMOVE>
__device__ __forceinline__ void decide_box(CBox& b, bool exceed) {
    BoxState* st = b.st;
    CGeom& g = geom_of(b);
    if (g.clear_last) st->last_valid = 0;
    if (st->state == 1u) {
        bool need;
        if (!st->last_valid || st->nlocal == 0) need = true;
        else if (g.every == 0) need = exceed;
        else if (g.check) need = st->steps_since_build >= g.every || exceed;
        else need = st->steps_since_build >= g.every;
        if (need) st->state = 0u;
        else st->steps_since_build += 1;
    }
    st->exceed = 0;
    st->final_pending = 0;  // applied by the kick-drift before this
    if (REMOVE) {
        if (st->state == 0u) st->nghost = 0;
        // The fused step also does sort_atoms_by_bin's bookkeeping here
        // (k_pre_sort): nothing before the sort changes its inputs. A
        // bounds change sets FullRebuild only in k_post_sort, after pbc's
        // run_if has seen the state decided here (neighbor.rs:837-841).
        st->do_sort = (st->nlocal > 0 && g.ncells > 0 && (g.periodic_sort || st->state == 0u || g.bounds_changed))
                          ? 1u
                          : 0u;
        st->not_identity = 0;
        st->local_cells_valid = 0;
        st->lb_fresh = 0;
        st->rebuild_now = st->state == 0u ? 1u : 0u;  // final value set by k_post_sort when it runs
        publish(b);
    }
}

template <bool REMOVE>
__global__ void k_decide(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        decide_box<REMOVE>(b, b.st->exceed != 0u);
    }
}

// classify_boundary_atoms (neighbor.rs:1321-1336): a local row with a ghost
// neighbour (index >= nlocal) is boundary; the flag goes to the slot scratch
__global__ void __launch_bounds__(BLK) k_classify(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const uint32_t nl = b.st->nlocal;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= nl) continue;
        uint32_t any = 0u;
        for (uint32_t e = b.offsets[i]; e < b.offsets[i + 1]; ++e) any |= b.indices[e] >= nl ? 1u : 0u;
        b.slot[i] = any;
    }
}

// compute_ke (atom.rs:589-604): each local's m (vx² + vy² + vz²) in the
// reference's order, summed per 256-atom tile in a fixed tree order into
// part[tile]; the tile sums are then added in tile order (k_sum_tiles or the
// host). `mask` (null: every local): one byte per local, non-zero = in the
// caller's group. Both buffers are scratch the caller provides.
__global__ void __launch_bounds__(BLK) k_ke(KARGS, const uint8_t* __restrict__ mask, double* __restrict__ part) {
    __shared__ double red[BLK / 64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const uint32_t i = tl.start + threadIdx.x;
        double v = 0.0;
        if (i < b.st->nlocal && (!mask || mask[i] != 0u)) {
            const double vx = b.vx[i], vy = b.vy[i], vz = b.vz[i];
            v = b.mass[i] * (vx * vx + vy * vy + vz * vz);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < BLK / 64; ++w) s += red[w];
            part[tl.start / BLK] = s;
        }
        __syncthreads();
    }
}

// *out = part[0] + part[1] + ... + part[n-1], added one after another from
// 0.0 in index order (the same order as a sequential host loop, so the sum is
// the same bits). One thread.
__global__ void k_sum_tiles(const double* __restrict__ part, uint32_t n, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (uint32_t k = 0; k < n; ++k) s += part[k];
    *out = s;
}

// counts[0..ncells] = 0 for the step's geometry: the shrink-wrap step stages
// its grid after the kick-drift, whose zeroing used the provisional grid
__global__ void __launch_bounds__(BLK) k_zero_counts(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const uint32_t c = tl.start + threadIdx.x;
        if (c <= geom_of(b).ncells) b.counts[c] = 0u;
    }
}

// shrink_wrap's extents (domain.rs:531-548): min and max of each coordinate
// over the locals, reduced as order-preserving u64 keys (negative values
// bit-inverted, positive ones with the sign bit set) so that atomic min/max
// give the exact extremes in any order
__device__ __forceinline__ unsigned long long ext_key(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ void k_extent_init(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        for (int k = 0; k < 3; ++k) {
            b.st->ext_key[k] = ~0ull;
            b.st->ext_key[3 + k] = 0ull;
        }
    }
}

__global__ void __launch_bounds__(BLK) k_extent(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;  // block-uniform
        const uint32_t i = tl.start + threadIdx.x;
        const bool act = i < b.st->nlocal;
        unsigned long long mn[3], mx[3];
        const double p[3] = {act ? b.x[i] : 0.0, act ? b.y[i] : 0.0, act ? b.z[i] : 0.0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = act ? ext_key(p[k]) : ~0ull;
            mx[k] = act ? ext_key(p[k]) : 0ull;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const unsigned long long a = __shfl_xor(mn[k], o, 64), c = __shfl_xor(mx[k], o, 64);
                mn[k] = a < mn[k] ? a : mn[k];
                mx[k] = c > mx[k] ? c : mx[k];
            }
        const bool any = __ballot(act) != 0ull;
        if ((threadIdx.x & 63u) == 0u && any) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(&b.st->ext_key[k], mn[k]);
                atomicMax(&b.st->ext_key[3 + k], mx[k]);
            }
        }
    }
}

// remove_ghost_atoms (atom.rs:625-629), FullRebuild only
__global__ void k_remove_ghosts(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& bb = box_at(boxes, tiles[t].box);
        if (!live(bb)) continue;
        BoxState* st = bb.st;
        if (st->state == 0u) st->nghost = 0;
    }
}

// run_if(in_state(FullRebuild)) for the per-phase entry points
__global__ void k_set_rebuild_now(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& bb = box_at(boxes, tiles[t].box);
        if (!live(bb)) continue;
        BoxState* st = bb.st;
        st->rebuild_now = st->state == 0u ? 1u : 0u;
    }
}

// zero_all_forces: atom.rs:631-635 (force + #[zero] columns of all atoms)
// zero_all_forces (atom.rs:631-635). In the fused step (CONSUMER_WRITES) a box
// whose consumer then assigns every force skips the force columns here.
template <bool CONSUMER_WRITES>
__global__ void __launch_bounds__(BLK) k_zero(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= b.st->nlocal + b.st->nghost) continue;
        if (!(CONSUMER_WRITES && b.consumer)) {
            b.fx[i] = 0.0;
            b.fy[i] = 0.0;
            b.fz[i] = 0.0;
        }
        for (int c = 0; c < b.ncols; ++c)
            if (b.cols[c].flags & 4u)
                for (int w = 0; w < b.cols[c].width; ++w) b.cols[c].data[(size_t)i * b.cols[c].width + w] = 0.0;
    }
}

// pbc / wrap_periodic / pbc_triclinic: domain.rs:594-729, one local's wrap
// (FullRebuild only; the caller checks the state and the index). Returns true
// when the atom left through a non-periodic face: pbc removes it
// (domain.rs:648-656, 700-706), which remove_atoms carries out for all such
// atoms at once.
__device__ __forceinline__ bool pbc_atom(CBox& b, uint32_t i) {
    CGeom& g = geom_of(b);
    if (g.triclinic) {
        double l[3];
        x2lamda(g, b.x[i], b.y[i], b.z[i], l[0], l[1], l[2]);
        int delta[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (g.periodic[k]) {
                if (l[k] < 0.0) { l[k] += 1.0; delta[k] = -1; }
                else if (l[k] >= 1.0) { l[k] -= 1.0; delta[k] = 1; }
            } else if (l[k] < 0.0 || l[k] >= 1.0) {
                return true;
            }
        }
        double x, y, z;
        lamda2x(g, l[0], l[1], l[2], x, y, z);
        b.x[i] = x;
        b.y[i] = y;
        b.z[i] = z;
        put_xr(b, i, x, y, z);
        b.imx[i] += delta[0];
        b.imy[i] += delta[1];
        b.imz[i] += delta[2];
        if (delta[1] != 0) {
            const double s = (double)delta[1];
            b.vx[i] -= s * g.bvel[0];
            b.vy[i] -= s * g.bvel[1];
            b.vz[i] -= s * g.bvel[2];
        }
        return false;
    }
    double* p[3] = {b.x, b.y, b.z};
    int32_t* im[3] = {b.imx, b.imy, b.imz};
    double q[3];
    bool moved = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = p[k][i];
        if (g.periodic[k]) {
            const double high = g.lo[k] + g.size[k];
            if (v < g.lo[k]) { v = v + g.size[k]; p[k][i] = v; im[k][i] -= 1; moved = true; }
            else if (v >= high) { v = v - g.size[k]; p[k][i] = v; im[k][i] += 1; moved = true; }
        } else if (v < g.lo[k] || v >= g.hi[k]) {
            return true;  // the axes after it are not wrapped: the atom goes
        }
        q[k] = v;
    }
    if (moved) put_xr(b, i, q[0], q[1], q[2]);  // xr already holds the unwrapped ones
    return false;
}

// pbc with removal. OPEN_ONLY (the fused step): only boxes with a
// non-periodic axis (fully periodic ones wrap inside the sort's count).
// A box that can remove writes one flag per local, and 0 at index nlocal,
// into the slot scratch, and counts its removals.
template <bool OPEN_ONLY>
__global__ void __launch_bounds__(BLK) k_pbc(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        if (st->state != 0u) continue;
        const bool open = !geom_of(b).pbc_all;
        if (OPEN_ONLY && !open) continue;
        const uint32_t i = tl.start + threadIdx.x;
        const uint32_t nl = st->nlocal;
        if (i > nl) continue;
        const bool rem = i < nl && pbc_atom(b, i);
        if (!open) continue;
        b.slot[i] = rem ? 1u : 0u;
        const uint64_t m = __ballot(rem);
        if (m && (uint32_t)__lane_id() == (uint32_t)__builtin_ctzll(m)) atomicAdd(&st->nremove, (uint32_t)__popcll(m));
    }
}

// Carries out the removals of this pbc (Atom::swap_remove over the locals in
// descending index order, atom.rs:268-271, 454-458, registry.swap_remove_all
// atom.rs:291-295 for the AtomData columns) in one parallel pass. With n
// locals of which m are removed, the k-th removal in that order moves the
// element then at position n - k into the removed slot, so the hole at a
// removed index p < n - m (its rank k = removals at or above p) ends up
// holding what position c = n - k holds when it is taken; when c was itself a
// removed position, that is what c received: c = n - rank(c), repeated. The
// chain rises strictly and ends at a kept atom in [n - m, n), so the copy
// source never lies below n - m and the copies into the holes do not overlap
// their sources. slot[] holds the exclusive scan of the removal flags
// (removals below each index), so rank(q) = m - slot[q] and flag(q) =
// slot[q + 1] - slot[q].
__device__ __forceinline__ void copy_atom(CBox& b, uint32_t s, uint32_t d) {
    b.x[d] = b.x[s]; b.y[d] = b.y[s]; b.z[d] = b.z[s];
    b.vx[d] = b.vx[s]; b.vy[d] = b.vy[s]; b.vz[d] = b.vz[s];
    b.fx[d] = b.fx[s]; b.fy[d] = b.fy[s]; b.fz[d] = b.fz[s];
    b.mass[d] = b.mass[s]; b.inv_mass[d] = b.inv_mass[s]; b.rad[d] = b.rad[s];
    b.tag[d] = b.tag[s]; b.type[d] = b.type[s]; b.origin[d] = b.origin[s];
    b.imx[d] = b.imx[s]; b.imy[d] = b.imy[s]; b.imz[d] = b.imz[s];
    const GMEM double4* xs = (const GMEM double4*)b.xr;
    ((GMEM double4*)b.xr)[d] = xs[s];
    b.lbx[d] = b.lbx[s]; b.lby[d] = b.lby[s]; b.lbz[d] = b.lbz[s];
    for (int c = 0; c < b.ncols; ++c) {
        const int w = b.cols[c].width;
        GMEM double* col = b.cols[c].data;
        for (int k = 0; k < w; ++k) col[(size_t)d * w + k] = col[(size_t)s * w + k];
    }
}

__global__ void __launch_bounds__(BLK) k_remove_fill(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        const uint32_t m = st->nremove;
        if (m == 0u || st->state != 0u) continue;
        const uint32_t n = st->nlocal, keep = n - m;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= keep || b.slot[i + 1] == b.slot[i]) continue;  // kept in place
        uint32_t c = n - (m - b.slot[i]);
        uint32_t hops = 0;  // at most m: the chain rises through removed positions
        while (b.slot[c + 1] != b.slot[c] && hops < m) {
            c = n - (m - b.slot[c]);
            ++hops;
        }
        if (c >= n || c < keep || b.slot[c + 1] != b.slot[c]) {  // not a kept tail atom: flag, copy nothing
            atomicOr(&b.st->err, ERR_REMOVE_PLAN);
            continue;
        }
        copy_atom(b, c, i);
    }
}

// nlocal drops by the removals; the state is already FullRebuild (pbc runs
// only then), so the step rebuilds its ghosts and lists for the new count
__global__ void k_remove_commit(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        const uint32_t m = st->nremove;
        if (m == 0u) continue;
        st->nlocal -= m;
        st->removed_total += m;
        st->nremove = 0u;
        st->last_valid = 0u;
        st->local_cells_valid = 0u;
    }
}

// sort_atoms_by_bin bookkeeping before the sort (one thread per box)
__global__ void k_pre_sort(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        CGeom& g = geom_of(b);
        if (g.bounds_changed) st->state = 0u;
        st->do_sort = (st->nlocal > 0 && g.ncells > 0 && (g.periodic_sort || st->state == 0u)) ? 1u : 0u;
        st->not_identity = 0;
        st->local_cells_valid = 0;
        st->lb_fresh = 0;
    }
}

// after the sort: a periodic sort that moved atoms forces a rebuild
// (neighbor.rs:989-996). FUSED also swaps the permuted column buffers and does
// the build's bookkeeping (bin_neighbor_list sets CommunicateOnly and resets the
// rebuild counters, neighbor.rs:1023-1024 / 669-690).
template <bool FUSED>
__global__ void k_post_sort(KARGS);

// ── stable counting sort by key ────────────────────────────────────────────
// SPATIAL: locals by cell (perm into `sorted`), BUILD: locals+ghosts by cell.
// BUILD with `fast`: the locals are already in cell order from this step's
// spatial sort (the reference reuses sorted_local_cell, neighbor.rs:1036-1064):
// `counts` still holds the local counts per cell and `sstart` their starts, so
// only the ghosts are counted, scattered and ranked; a local k goes to
// bin_start[c] + (k - sstart[c]).
template <int W>
struct SortView {
    uint32_t n, nkeys;
    bool active, fast;
    uint32_t *key, *slot, *counts, *start, *tmp, *out;
};

template <int W>
__device__ __forceinline__ SortView<W> sort_view(CBox& b) {
    SortView<W> v;
    const BoxState* st = b.st;
    CGeom& g = geom_of(b);
    v.fast = false;
    if (W == SORT_SPATIAL) {
        v.n = st->nlocal; v.nkeys = g.ncells; v.active = st->do_sort != 0;
        v.key = b.scell; v.slot = b.slot; v.counts = b.counts; v.start = b.sstart; v.tmp = b.tmp; v.out = b.perm;
    } else if (W == SORT_BUILD) {
        v.n = st->nlocal + st->nghost; v.nkeys = g.ncells;
        v.active = st->rebuild_now != 0 && (st->err & (ERR_GHOST_CAPACITY)) == 0;
        v.fast = st->local_cells_valid != 0;
        v.key = b.cell; v.slot = b.slot; v.counts = b.counts; v.start = b.bin_start; v.tmp = b.tmp; v.out = b.sorted;
    }
    return v;
}

template <int W>
__global__ void __launch_bounds__(BLK) k_sort_zero(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        SortView<W> v = sort_view<W>(b);
        if (!v.active || v.fast) continue;
        const uint32_t c = tl.start + threadIdx.x;
        if (c <= v.nkeys) v.counts[c] = 0u;
    }
}

// PBC (the fused step's spatial sort): each thread first wraps its local, as
// k_pbc would; the count reads only the thread's own atom.
template <int W, bool PBC = false>
__global__ void __launch_bounds__(BLK) k_sort_count(KARGS) {
    FOR_TILES_SPREAD_IF(t, W == SORT_BUILD) {  // the build counts only ghosts when it reuses the sort
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        SortView<W> v = sort_view<W>(b);
        const uint32_t i = tl.start + threadIdx.x;
        if (PBC && b.st->state == 0u && i < b.st->nlocal && geom_of(b).pbc_all) (void)pbc_atom(b, i);
        if (!v.active) continue;
        // Cell sorts: the input is nearly cell-ordered (locals were sorted at the
        // last rebuild, ghosts follow their sources), so lanes form runs of equal
        // keys; one atomic per run hands out consecutive slots.
        const uint32_t first = v.fast ? b.st->nlocal : 0u;
        if (v.fast && i < first) v.key[i] = b.scell[b.perm[i]];  // sorted_local_cell
        const bool act = i >= first && i < v.n;
        uint32_t key = 0xffffffffu;
        if (act) {
            key = cell_of(geom_of(b), b.x[i], b.y[i], b.z[i]);
            v.key[i] = key;
        }
        const uint32_t lane = __lane_id();
        const uint32_t prev = __shfl_up(key, 1, 64);
        const bool head = act && (lane == 0u || prev != key);
        const uint64_t heads = __ballot(head);
        const uint64_t actm = __ballot(act);
        const uint32_t nact = actm ? 64u - (uint32_t)__builtin_clzll(actm) : 0u;  // one past the last active lane
        const uint64_t upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
        const uint32_t leader = 63u - (uint32_t)__builtin_clzll((heads & upto) | 1ull);
        uint32_t base = 0;
        if (head) {
            const uint64_t above = heads & ~((2ull << lane) - 1ull);
            const uint32_t next = (lane == 63u || above == 0ull) ? nact : (uint32_t)__builtin_ctzll(above);
            base = atomicAdd(&v.counts[key], next - lane);
        }
        base = __shfl(base, leader, 64);
        if (act) v.slot[i] = base + (lane - leader);
    }
}

template <int W>
__global__ void __launch_bounds__(BLK) k_sort_scatter(KARGS) {
    FOR_TILES_SPREAD_IF(t, W == SORT_BUILD) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        SortView<W> v = sort_view<W>(b);
        if (!v.active) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= v.n || (v.fast && i < b.st->nlocal)) continue;
        const uint32_t key = v.key[i];
        v.tmp[v.start[key] + v.slot[i]] = i;
    }
}

template <int W>
__global__ void __launch_bounds__(BLK) k_sort_rank(KARGS) {
    FOR_TILES_SPREAD_IF(t, W == SORT_BUILD) {  // ghosts rank against their cells, locals only copy
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        SortView<W> v = sort_view<W>(b);
        if (!v.active) continue;
        const uint32_t i = tl.start + threadIdx.x;
        const bool act = i < v.n;
        const uint32_t key = act ? v.key[i] : 0xffffffffu;
        // the wave's runs of equal keys (lanes), as k_sort_count forms them
        const uint32_t lane = __lane_id();
        const uint32_t prev = __shfl_up(key, 1, 64);
        const uint64_t heads = __ballot(act && (lane == 0u || prev != key));
        const uint64_t actm = __ballot(act);
        if (!act) continue;
        const uint32_t nact = 64u - (uint32_t)__builtin_clzll(actm);
        const uint64_t upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
        const uint32_t leader = 63u - (uint32_t)__builtin_clzll((heads & upto) | 1ull);
        const uint64_t above = heads & ~upto;
        const uint32_t run_end = above ? (uint32_t)__builtin_ctzll(above) : nact;
        uint32_t s = v.start[key], e = v.start[key + 1];
        uint32_t dst;
        if (!v.fast && e - s == run_end - leader) {
            // the run is the whole cell: its atoms are consecutive indices, in order
            dst = s + (lane - leader);
        } else if (v.fast) {
            const uint32_t ls = b.sstart[key], nloc = b.sstart[key + 1] - ls;
            if (i < b.st->nlocal) {
                dst = s + (i - ls);
            } else {
                s += nloc;  // the cell's ghosts follow its locals
                uint32_t r = 0;
                if (e - s > 1u)
                    for (uint32_t m = s; m < e; ++m) r += v.tmp[m] < i ? 1u : 0u;
                dst = s + r;
            }
        } else {
            uint32_t r = 0;
            if (e - s > 1u)
                for (uint32_t m = s; m < e; ++m) r += v.tmp[m] < i ? 1u : 0u;
            dst = s + r;
        }
        v.out[dst] = i;
        if (W == SORT_BUILD) {
            b.sidx[i] = dst;
            ((double4*)b.spos)[dst] = ((const double4*)b.xr)[i];  // bin_sorted_pos (+ radius)
        }
        if (W == SORT_SPATIAL) {
            wave_flag(&b.st->not_identity, dst != i);
        }
    }
}

// atom.rs:479-490 apply_permutation on every permuted column (gather into the
// alternates; forces and #[zero] columns are zero at this point of the step,
// and a local's origin_index is always 0 on one rank, so those stay in place)
// FUSED (the step's sort, always followed by this step's build): also writes
// the permuted positions as the build's last_build_pos (save_build_positions,
// neighbor.rs:669-690), which k_save_build then skips.
template <bool FUSED>
__global__ void __launch_bounds__(BLK) k_permute(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->do_sort || !st->not_identity) continue;
        const uint32_t k = tl.start + threadIdx.x;
        if (k >= st->nlocal) continue;
        const uint32_t p = b.perm[k];
        const double px = b.x[p], py = b.y[p], pz = b.z[p], pr = b.rad[p];
        b.ax[k] = px;
        b.ay[k] = py;
        b.az[k] = pz;
        if (FUSED) {
            b.lbx[k] = px;
            b.lby[k] = py;
            b.lbz[k] = pz;
        }
        {
            double* q = b.xr + 4 * (size_t)k;
            q[0] = px;
            q[1] = py;
            q[2] = pz;
            q[3] = pr;
        }
        b.avx[k] = b.vx[p];
        b.avy[k] = b.vy[p];
        b.avz[k] = b.vz[p];
        b.amass[k] = b.mass[p];
        b.ainv_mass[k] = b.inv_mass[p];
        b.arad[k] = pr;
        b.atag[k] = b.tag[p];
        b.atype[k] = b.type[p];
        b.aimx[k] = b.imx[p];
        b.aimy[k] = b.imy[p];
        b.aimz[k] = b.imz[p];
        for (int c = 0; c < b.ncols; ++c) {
            const ColDev& cd = b.cols[c];
            if (cd.flags & 4u) continue;
            for (int w = 0; w < cd.width; ++w) cd.alt[(size_t)k * cd.width + w] = cd.data[(size_t)p * cd.width + w];
        }
    }
}

template <typename T>
__device__ __forceinline__ void swap_ptr(T*& a, T*& b) {
    T* t = a;
    a = b;
    b = t;
}

__global__ void k_swap_ptrs(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        BoxDev& b = *boxes[tiles[t].box];
        if (!live(b)) continue;
        if (!b.st->do_sort || !b.st->not_identity) continue;
        swap_ptr(b.x, b.ax); swap_ptr(b.y, b.ay); swap_ptr(b.z, b.az);
        swap_ptr(b.vx, b.avx); swap_ptr(b.vy, b.avy); swap_ptr(b.vz, b.avz);
        swap_ptr(b.mass, b.amass); swap_ptr(b.inv_mass, b.ainv_mass); swap_ptr(b.rad, b.arad);
        swap_ptr(b.tag, b.atag); swap_ptr(b.type, b.atype);
        swap_ptr(b.imx, b.aimx); swap_ptr(b.imy, b.aimy); swap_ptr(b.imz, b.aimz);
        for (int c = 0; c < b.ncols; ++c)
            if (!(b.cols[c].flags & 4u)) swap_ptr(b.cols[c].data, b.cols[c].alt);
    }
}

template <bool FUSED>
__device__ __forceinline__ void post_sort_box(BoxDev& b) {
    BoxState* st = b.st;
    CGeom& g = geom_of(b);
    if (g.bounds_changed) st->state = 0u;
    if (FUSED && st->do_sort && st->not_identity) {
        swap_ptr(b.x, b.ax); swap_ptr(b.y, b.ay); swap_ptr(b.z, b.az);
        swap_ptr(b.vx, b.avx); swap_ptr(b.vy, b.avy); swap_ptr(b.vz, b.avz);
        swap_ptr(b.mass, b.amass); swap_ptr(b.inv_mass, b.ainv_mass); swap_ptr(b.rad, b.arad);
        swap_ptr(b.tag, b.atag); swap_ptr(b.type, b.atype);
        swap_ptr(b.imx, b.aimx); swap_ptr(b.imy, b.aimy); swap_ptr(b.imz, b.aimz);
        for (int c = 0; c < b.ncols; ++c)
            if (!(b.cols[c].flags & 4u)) swap_ptr(b.cols[c].data, b.cols[c].alt);
    }
    if (g.periodic_sort && st->do_sort && st->not_identity) {
        st->last_valid = 0;
        st->state = 0u;
    }
    if (st->do_sort) st->local_cells_valid = 1;
    if (FUSED && st->do_sort && st->not_identity) st->lb_fresh = 1;
    st->rebuild_now = st->state == 0u ? 1u : 0u;
    if (FUSED && st->rebuild_now) {
        st->state = 1u;
        st->steps_since_build = 0;
        st->last_valid = 1;
        st->rebuilds += 1;
    }
}

template <bool FUSED>
__global__ void k_post_sort(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        BoxDev& b = *boxes[tiles[t].box];
        if (!live(b)) continue;
        post_sort_box<FUSED>(b);
    }
}

// the incoming-pair transpose is rebuilt with the CSR, for boxes whose
// consumer or force hook reads it
__device__ __forceinline__ bool transpose_active(CBox& b) {
    const BoxState* st = b.st;
    return st->rebuild_now != 0 && b.transpose && (st->err & (ERR_PAIR_CAPACITY | ERR_GHOST_CAPACITY)) == 0;
}
// the transpose's two strategies: arrival slots from atomics in the CSR fill or
// placement, a scan and a scatter (k_tplace); or, for the default build of a
// box with tpull set, the atomic-free pull (k_tpull_count, k_tpull_place). The
// count of the default build latches the strategy for the rest of its build,
// so a descriptor written while the build is queued cannot split it.
__device__ __forceinline__ bool tpull_on(CBox& b) { return b.tpull != 0 && b.csr_mode == 5; }
__device__ __forceinline__ bool tpull_build(CBox& b) { return b.csr_mode == 5 && b.st->tpull_now != 0u; }
__device__ __forceinline__ bool tatomic_active(CBox& b) { return transpose_active(b) && !tpull_build(b); }
__device__ __forceinline__ bool tpull_active(CBox& b) { return transpose_active(b) && tpull_build(b); }

// ── exclusive scans (reduce tiles -> scan tile sums per box -> down-sweep) ──
template <int W>
__device__ __forceinline__ void scan_view(CBox& b, uint32_t& n, bool& active, GMEM uint32_t*& in, GMEM uint32_t*& out) {
    const BoxState* st = b.st;
    if (W == SCAN_CELLS_SPATIAL || W == SCAN_CELLS_BUILD) {
        n = geom_of(b).ncells + 1;
        active = W == SCAN_CELLS_SPATIAL ? st->do_sort != 0
                                         : (st->rebuild_now != 0 && (st->err & ERR_GHOST_CAPACITY) == 0);
        in = b.counts;
        out = W == SCAN_CELLS_SPATIAL ? b.sstart : b.bin_start;
    } else if (W == SCAN_CSR) {
        n = st->nlocal + 1;
        active = st->rebuild_now != 0 && (st->err & ERR_GHOST_CAPACITY) == 0;
        in = b.offsets;
        out = b.offsets;
    } else if (W == SCAN_REMOVE) {
        n = st->nlocal + 1;
        active = st->nremove != 0 && st->state == 0u;
        in = b.slot;
        out = b.slot;
    } else if (W == SCAN_WAVES) {
        n = ((st->nlocal + 63u) >> 6) + 1u;
        active = st->rebuild_now != 0 && (st->err & ERR_GHOST_CAPACITY) == 0;
        in = b.wcount;
        out = b.wcount;
    } else {
        n = st->nlocal + st->nghost + 1;
        active = tatomic_active(b);
        in = b.tcount;
        out = b.tstart;
    }
}

template <int W>
__global__ void __launch_bounds__(BLK) k_scan_reduce(KARGS) {
    __shared__ uint32_t lds[BLK / 64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        uint32_t n;
        bool active;
        GMEM uint32_t *in, *out;
        scan_view<W>(b, n, active, in, out);
        if (!active || tl.start >= n) continue;  // block-uniform
        uint32_t s = 0;
        const uint32_t base = tl.start + threadIdx.x * SCAN_PER_THREAD;
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k)
            if (base + k < n) s += in[base + k];
        uint32_t tot;
        block_excl_scan(s, lds, tot);
        if (threadIdx.x == 0) b.scan_part[tl.start / SCAN_TILE] = tot;
    }
}

// Down-sweep: each tile adds the totals of the box's earlier tiles (summed
// here from k_scan_reduce's per-tile totals, a few loads per thread) to its
// own exclusive scan; the first tile also sums them all for the CSR count.
// PUB (the plain step's CSR offsets): the box's first tile also publishes the
// build's state to the host mirror, active or not.
template <int W, bool PUB = false>
__global__ void __launch_bounds__(BLK) k_scan_down(KARGS) {
    __shared__ uint32_t lds[BLK / 64];
    __shared__ uint32_t red[BLK / 64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        uint32_t n;
        bool active;
        GMEM uint32_t *in, *out;
        scan_view<W>(b, n, active, in, out);
        if (!active || tl.start >= n) {  // block-uniform
            if (PUB && tl.start == 0u && threadIdx.x == 0)
                publish_vals(b, __hip_atomic_load(&b.st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                             b.st->npairs);
            continue;
        }
        const uint32_t tix = tl.start / SCAN_TILE;
        const bool first = tix == 0u;
        // the tile's own entries first: their loads are in flight while the
        // earlier tiles' totals are summed
        uint32_t v[SCAN_PER_THREAD];
        const uint32_t base = tl.start + threadIdx.x * SCAN_PER_THREAD;
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k) v[k] = base + k < n ? in[base + k] : 0u;
        const uint32_t lim = first ? (n + SCAN_TILE - 1) / SCAN_TILE : tix;
        uint32_t acc = 0;
        for (uint32_t k = threadIdx.x; k < lim; k += BLK) acc += b.scan_part[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        acc = 0;
#pragma unroll
        for (uint32_t w = 0; w < BLK / 64; ++w) acc += red[w];
        __syncthreads();  // red[] is reused by the next tile
        // the tile-staged CSR build's group sums start from zero (its count adds
        // into them; this scan runs before it in every build)
        if (W == SCAN_CELLS_BUILD && first)
            for (uint32_t q = threadIdx.x; q < b.cap_local / (BLK * CSR1_GROUP) + 2u; q += BLK) b.cgroup[q] = 0u;
        if ((W == SCAN_CSR || W == SCAN_WAVES) && first && threadIdx.x == 0) {
            BoxState* st = b.st;
            st->local_cells_valid = 0;  // consumed by this build (neighbor.rs:596)
            st->npairs = acc;
            uint32_t err;  // the bits as they stand after this kernel's own update
            if (acc > b.cap_pairs) {
                err = atomicOr(&st->err, ERR_PAIR_CAPACITY) | ERR_PAIR_CAPACITY;
            } else {
                err = __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                CGeom& g = geom_of(b);
                const uint64_t na = (uint64_t)st->nlocal + st->nghost;
                const uint64_t per = 24u + (g.uniform ? 0u : 8u) + 4u;
                st->csr_bytes += per * na + 4u * ((uint64_t)g.ncells + 1u) + 4u * ((uint64_t)st->nlocal + 1u) +
                                 4u * (uint64_t)acc;
                st->csr_builds += 1;
            }
            if (PUB) publish_vals(b, err, acc);
        }
        const uint32_t prefix = first ? 0u : acc;
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k) sum += v[k];
        uint32_t tot;
        uint32_t pre = block_excl_scan(sum, lds, tot) + prefix;
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k) {
            if (base + k < n) out[base + k] = pre;
            pre += v[k];
        }
    }
}

// ── borders: comm.rs:682-860 + pack_border_atoms :288-358 ──────────────────
// selection value of a position along dim d: the coordinate, or lamda when triclinic
__device__ __forceinline__ double border_coord(CGeom& g, double x, double y, double z, int d) {
    if (g.triclinic) {
        double lx, ly, lz;
        x2lamda(g, x, y, z, lx, ly, lz);
        return d == 0 ? lx : (d == 1 ? ly : lz);
    }
    return d == 0 ? x : (d == 1 ? y : z);
}

// Border tiles: BORDER_TILE consecutive atoms per block; wave w reads
// BORDER_PER chunks of 64 consecutive atoms (coalesced), chunk c of wave w
// starting at tile + (w * BORDER_PER + c) * 64.
__device__ __forceinline__ uint32_t border_atom(uint32_t tile_start, uint32_t wid, int c, uint32_t lane) {
    return tile_start + (wid * BORDER_PER + (uint32_t)c) * 64u + lane;
}

// Selection counts per border tile for all three dims at once: bcnt[6 t + 2 d
// + side] = atoms of tile t among the locals that swap (d, side) selects.
// Tiles past the locals start at zero; each dim's write adds the ghosts it
// creates to the later dims' counts (their scans include those ghosts).
__global__ void __launch_bounds__(BLK) k_border_count(KARGS) {
    __shared__ uint32_t wc[6][BLK / 64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        CGeom& g = geom_of(b);
        if (!st->rebuild_now) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t tix = tl.start / BORDER_TILE;
        if (tl.start >= nl) {  // block-uniform
            if (threadIdx.x < 6u) b.bcnt[6 * tix + threadIdx.x] = 0u;
            continue;
        }
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        uint32_t c[6] = {0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < BORDER_PER; ++k) {
            const uint32_t i = border_atom(tl.start, wid, k, lane);
            const bool in = i < nl;
            double x = 0.0, y = 0.0, z = 0.0;
            if (in) {
                x = b.x[i];
                y = b.y[i];
                z = b.z[i];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                bool p0 = false, p1 = false;
                if (in) {  // non-periodic swaps select too (send_indices) but receive nothing
                    const double v = border_coord(g, x, y, z, d);
                    p0 = v < g.thr_lo[d];
                    p1 = v >= g.thr_hi[d];
                }
                c[2 * d] += (uint32_t)__popcll(__ballot(p0));
                c[2 * d + 1] += (uint32_t)__popcll(__ballot(p1));
            }
        }
        if (lane == 0u)
#pragma unroll
            for (int q = 0; q < 6; ++q) wc[q][wid] = c[q];
        __syncthreads();
        if (threadIdx.x < 6u) {
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t w = 0; w < BLK / 64; ++w) sum += wc[threadIdx.x][w];
            b.bcnt[6 * tix + threadIdx.x] = sum;
        }
        __syncthreads();
    }
}

// pack_border + unpack_atom (atom.rs:525-571), split in two: the dim's kernel
// writes what later dims and the chain need (position = source + shift, the
// source index and the swap); k_ghost_fill copies the rest once all six swaps
// are placed.
__device__ __forceinline__ void place_ghost(CBox& b, CGeom& g, uint32_t dst, uint32_t i, int s, double& gx,
                                            double& gy, double& gz) {
    // the source's {x, y, z, rad} record: for a ghost source that is the one
    // placed by an earlier dim (its other fields are filled later)
    const double4 p = ((const double4*)b.xr)[i];
    gx = p.x + g.swap_shift[s][0];
    gy = p.y + g.swap_shift[s][1];
    gz = p.z + g.swap_shift[s][2];
    b.x[dst] = gx;
    b.y[dst] = gy;
    b.z[dst] = gz;
    ((double4*)b.xr)[dst] = make_double4(gx, gy, gz, p.w);
    b.origin[dst] = (int32_t)i;
    b.gswap[dst] = (uint8_t)s;
}

// The remaining ghost fields. A ghost of a ghost copies its source, which
// copied its own source, so every copied field equals the local at the root
// of the chain; the velocity adds each swap's shift root-first, as the
// swap-by-swap copies do; inv_mass = 1 / mass as in unpack_atom.
__device__ __forceinline__ void fill_ghost(CBox& b, const BoxState* st, uint32_t gi) {
    const uint32_t nl = st->nlocal;
    CGeom& g = geom_of(b);
    int sw[3];
    int depth = 0;
    uint32_t r = gi;
    while (r >= nl && depth < 3) {
        sw[depth++] = b.gswap[r];
        r = (uint32_t)b.origin[r];
    }
    double qx = b.vx[r], qy = b.vy[r], qz = b.vz[r];
    for (int k = depth - 1; k >= 0; --k) {
        const int s = sw[k];
        qx = qx + g.swap_vshift[s][0];
        qy = qy + g.swap_vshift[s][1];
        qz = qz + g.swap_vshift[s][2];
    }
    b.vx[gi] = qx;
    b.vy[gi] = qy;
    b.vz[gi] = qz;
    b.fx[gi] = b.fx[r];
    b.fy[gi] = b.fy[r];
    b.fz[gi] = b.fz[r];
    const double m = b.mass[r];
    b.mass[gi] = m;
    b.inv_mass[gi] = 1.0 / m;
    b.rad[gi] = b.rad[r];
    b.tag[gi] = b.tag[r];
    b.type[gi] = b.type[r];
    b.imx[gi] = b.imx[r];
    b.imy[gi] = b.imy[r];
    b.imz[gi] = b.imz[r];
    for (int c = 0; c < b.ncols; ++c) {
        const ColDev& cd = b.cols[c];
        for (int w = 0; w < cd.width; ++w) cd.data[(size_t)gi * cd.width + w] = cd.data[(size_t)r * cd.width + w];
    }
}

__global__ void __launch_bounds__(BLK) k_ghost_fill(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t gi = nl + tl.start + threadIdx.x;
        if (gi >= nl + st->nghost) continue;
        fill_ghost(b, st, gi);
    }
}

// Dim d of borders (comm.rs:736-853) in one launch. Every block sums the
// dim's per-tile counts (bcnt) for the totals and for its own tile's prefix,
// lists its tile's selected atoms in LDS in scan order (lo swap, then hi swap)
// and places one ghost per lane from the list: a tile's ghosts are
// consecutive, so the writes coalesce. The first tile's block writes the dim's
// swap table. Ghosts created here are counted (LDS, then one atomic per tile)
// into the selections of the later dims, whose scans include them.
__global__ void __launch_bounds__(BLK) k_border_write(KARGS, int d) {
    __shared__ uint32_t wc[2][BLK / 64];
    __shared__ uint32_t red[4][BLK / 64];
    __shared__ uint32_t list[2 * BORDER_TILE];
    __shared__ uint32_t lc[16];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        CGeom& g = geom_of(b);
        if (!st->rebuild_now) continue;
        const uint32_t nl = st->nlocal;
        // candidates: locals and the ghosts of the earlier dims (borders truncates old ghosts first)
        const uint32_t scan_end = d == 0 ? nl : st->swap_start[2 * d - 1] + st->swap_count[2 * d - 1];
        const uint32_t tix = tl.start / BORDER_TILE;
        if (st->err & ERR_GHOST_CAPACITY) continue;
        if (!g.periodic[d] && tix != 0) continue;  // only the totals are needed there (block-uniform)
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        // the tile's coordinates first (lamda needs all three): their loads are
        // in flight while the dim's counts are summed
        const bool places = g.periodic[d] && tl.start < scan_end;  // block-uniform
        double cx[BORDER_PER], cy[BORDER_PER], cz[BORDER_PER];
#pragma unroll
        for (int c = 0; c < BORDER_PER; ++c) {
            const uint32_t i = border_atom(tl.start, wid, c, lane);
            cx[c] = cy[c] = cz[c] = 0.0;
            if (places && i < scan_end) {
                if (g.triclinic) {
                    cx[c] = b.x[i];
                    cy[c] = b.y[i];
                    cz[c] = b.z[i];
                } else {
                    cx[c] = d == 0 ? b.x[i] : (d == 1 ? b.y[i] : b.z[i]);
                }
            }
        }
        // totals of the dim and this tile's exclusive prefixes
        const uint32_t nt = (scan_end + BORDER_TILE - 1) / BORDER_TILE;
        uint32_t pb0 = 0, pb1 = 0, tot0 = 0, tot1 = 0;
        for (uint32_t k = threadIdx.x; k < nt; k += BLK) {
            const uint32_t v0 = b.bcnt[6 * k + 2 * d], v1 = b.bcnt[6 * k + 2 * d + 1];
            tot0 += v0;
            tot1 += v1;
            if (k < tix) {
                pb0 += v0;
                pb1 += v1;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            pb0 += __shfl_xor(pb0, o, 64);
            pb1 += __shfl_xor(pb1, o, 64);
            tot0 += __shfl_xor(tot0, o, 64);
            tot1 += __shfl_xor(tot1, o, 64);
        }
        if (lane == 0u) {
            red[0][wid] = pb0;
            red[1][wid] = pb1;
            red[2][wid] = tot0;
            red[3][wid] = tot1;
        }
        if (threadIdx.x < 16u) lc[threadIdx.x] = 0u;
        __syncthreads();
        pb0 = pb1 = tot0 = tot1 = 0;
#pragma unroll
        for (uint32_t w = 0; w < BLK / 64; ++w) {
            pb0 += red[0][w];
            pb1 += red[1][w];
            tot0 += red[2][w];
            tot1 += red[3][w];
        }
        if (!g.periodic[d]) {  // no image (to_proc = -1 on one rank): the selection is only recorded
            if (tix == 0 && threadIdx.x == 0) {
                st->swap_start[2 * d] = scan_end;
                st->swap_count[2 * d] = 0u;
                st->swap_send[2 * d] = tot0;
                st->swap_start[2 * d + 1] = scan_end;
                st->swap_count[2 * d + 1] = 0u;
                st->swap_send[2 * d + 1] = tot1;
                st->nghost = scan_end - nl;
            }
            __syncthreads();  // red[] is reused by the next tile
            continue;
        }
        const bool over = (uint64_t)scan_end + tot0 + tot1 > (uint64_t)b.cap_atoms;
        if (tix == 0 && threadIdx.x == 0) {
            if (over) {
                atomicOr(&st->err, ERR_GHOST_CAPACITY);
            } else {
                st->swap_start[2 * d] = scan_end;
                st->swap_count[2 * d] = tot0;
                st->swap_send[2 * d] = tot0;
                st->swap_start[2 * d + 1] = scan_end + tot0;
                st->swap_count[2 * d + 1] = tot1;
                st->swap_send[2 * d + 1] = tot1;
                st->nghost = scan_end + tot0 + tot1 - nl;
            }
        }
        __syncthreads();  // red[] is reused by the next tile
        if (over || tl.start >= scan_end) continue;  // block-uniform
        uint64_t m0[BORDER_PER], m1[BORDER_PER];
        uint32_t c0 = 0, c1 = 0;
#pragma unroll
        for (int c = 0; c < BORDER_PER; ++c) {
            const uint32_t i = border_atom(tl.start, wid, c, lane);
            bool p0 = false, p1 = false;
            if (i < scan_end) {
                const double v = g.triclinic ? border_coord(g, cx[c], cy[c], cz[c], d) : cx[c];
                p0 = v < g.thr_lo[d];
                p1 = v >= g.thr_hi[d];
            }
            m0[c] = __ballot(p0);
            m1[c] = __ballot(p1);
            c0 += (uint32_t)__popcll(m0[c]);
            c1 += (uint32_t)__popcll(m1[c]);
        }
        if (lane == 0u) {
            wc[0][wid] = c0;
            wc[1][wid] = c1;
        }
        __syncthreads();
        uint32_t w0 = 0, w1 = 0, t0 = 0, t1 = 0;
#pragma unroll
        for (uint32_t w = 0; w < BLK / 64; ++w) {
            if (w < wid) {
                w0 += wc[0][w];
                w1 += wc[1][w];
            }
            t0 += wc[0][w];
            t1 += wc[1][w];
        }
#pragma unroll
        for (int c = 0; c < BORDER_PER; ++c) {
            const uint32_t i = border_atom(tl.start, wid, c, lane);
            if ((m0[c] >> lane) & 1ull) list[w0 + lane_prefix(m0[c])] = i;
            if ((m1[c] >> lane) & 1ull) list[t0 + w1 + lane_prefix(m1[c])] = i;
            w0 += (uint32_t)__popcll(m0[c]);
            w1 += (uint32_t)__popcll(m1[c]);
        }
        __syncthreads();
        const uint32_t base0 = scan_end + pb0;
        const uint32_t base1 = scan_end + tot0 + pb1;
        // a tile's lo (hi) ghosts span at most two border tiles: lc[range*8 + which*4 + later*2 + side]
        const uint32_t ft0 = base0 / BORDER_TILE, ft1 = base1 / BORDER_TILE;
        for (uint32_t r = threadIdx.x; r < t0 + t1; r += BLK) {
            const uint32_t i = list[r];
            const bool lo = r < t0;
            const uint32_t dst = lo ? base0 + r : base1 + (r - t0);
            double gx, gy, gz;
            place_ghost(b, g, dst, i, lo ? 2 * d : 2 * d + 1, gx, gy, gz);
            for (int e = d + 1; e < 3; ++e) {
                const double v = border_coord(g, gx, gy, gz, e);
                const uint32_t which = dst / BORDER_TILE - (lo ? ft0 : ft1);
                const uint32_t slot = (lo ? 0u : 8u) + which * 4u + (uint32_t)(e - d - 1) * 2u;
                if (v < g.thr_lo[e]) atomicAdd(&lc[slot], 1u);
                if (v >= g.thr_hi[e]) atomicAdd(&lc[slot + 1u], 1u);
            }
        }
        __syncthreads();
        if (d < 2 && threadIdx.x < 16u && lc[threadIdx.x]) {
            const uint32_t q = threadIdx.x;
            const uint32_t tile = ((q & 8u) ? ft1 : ft0) + ((q >> 2) & 1u);
            const uint32_t e = (uint32_t)d + 1u + ((q >> 1) & 1u);
            atomicAdd(&b.bcnt[6 * tile + 2 * e + (q & 1u)], lc[q]);
        }
        __syncthreads();
    }
}

// ── forward comm: comm.rs:400-432, 496-592 ─────────────────────────────────
// A ghost refreshes from the chain ghost -> ... -> local root, adding the
// fixed swap offsets root-first, which equals the reference's swap-by-swap order.
__device__ __forceinline__ void forward_ghost(CBox& b, const BoxState* st, uint32_t gi) {
    const uint32_t nl = st->nlocal;
    CGeom& g = geom_of(b);
    int sw[3];
    int depth = 0;
    uint32_t cur = gi;
    while (cur >= nl && depth < 3) {
        sw[depth++] = b.gswap[cur];
        cur = (uint32_t)b.origin[cur];
    }
    double px = b.x[cur], py = b.y[cur], pz = b.z[cur];
    double qx = b.vx[cur], qy = b.vy[cur], qz = b.vz[cur];
    for (int k = depth - 1; k >= 0; --k) {
        const int s = sw[k];
        px = px + g.swap_shift[s][0];
        py = py + g.swap_shift[s][1];
        pz = pz + g.swap_shift[s][2];
        qx = qx + g.swap_vshift[s][0];
        qy = qy + g.swap_vshift[s][1];
        qz = qz + g.swap_vshift[s][2];
    }
    b.x[gi] = px;
    b.y[gi] = py;
    b.z[gi] = pz;
    put_xr(b, gi, px, py, pz);
    b.vx[gi] = qx;
    b.vy[gi] = qy;
    b.vz[gi] = qz;
    for (int c = 0; c < b.ncols; ++c) {
        const ColDev& cd = b.cols[c];
        if (!(cd.flags & 1u)) continue;
        for (int w = 0; w < cd.width; ++w) cd.data[(size_t)gi * cd.width + w] = cd.data[(size_t)cur * cd.width + w];
    }
}

__global__ void __launch_bounds__(BLK) k_forward(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (st->rebuild_now) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t gi = nl + tl.start + threadIdx.x;
        if (gi >= nl + st->nghost) continue;
        forward_ghost(b, st, gi);
    }
}

// ── reverse comm: comm.rs:867-878, 938-1024, one launch per swap ───────────
// Round of dimension d. The lo swap's sources (x < lo+gc) and the hi swap's
// (x >= hi-gc) are different atoms when the two selections cannot overlap
// (Geom::disjoint), and a dim's sources all precede its own ghosts, so both
// swaps of the dim then add onto distinct owners and run in one launch
// (part 0). Otherwise part 0 is the lo swap and part 1 the hi swap, in the
// reference order.
__global__ void __launch_bounds__(BLK) k_reverse(KARGS, int d, int part) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        CGeom& g = geom_of(b);
        // a failed build is redone by the host, and this reverse pass with it
        if (!g.newton || (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY))) continue;
        const bool both = g.disjoint[d] != 0;
        uint32_t k = tl.start + threadIdx.x;
        int s;
        if (part == 0) {
            const uint32_t c0 = st->swap_count[2 * d];
            const uint32_t n = both ? c0 + st->swap_count[2 * d + 1] : c0;
            if (k >= n) continue;
            s = k < c0 ? 2 * d : 2 * d + 1;
            if (k >= c0) k -= c0;
        } else {
            if (both) continue;
            s = 2 * d + 1;
            if (k >= st->swap_count[s]) continue;
        }
        const uint32_t gi = st->swap_start[s] + k;
        const uint32_t o = (uint32_t)b.origin[gi];
        b.fx[o] += b.fx[gi];
        b.fy[o] += b.fy[gi];
        b.fz[o] += b.fz[gi];
        for (int c = 0; c < b.ncols; ++c) {
            const ColDev& cd = b.cols[c];
            if (!(cd.flags & 2u)) continue;
            for (int w = 0; w < cd.width; ++w) cd.data[(size_t)o * cd.width + w] += cd.data[(size_t)gi * cd.width + w];
        }
    }
}

// ── neighbor build: neighbor.rs:1013-1319 ──────────────────────────────────
__global__ void k_build_prep(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        if (!st->rebuild_now) continue;
        st->state = 1u;  // CommunicateOnly
        st->steps_since_build = 0;
        st->last_valid = 1;
        st->rebuilds += 1;
    }
}

// save_build_positions (neighbor.rs:669-690); radius extrema are fixed per box
__device__ __forceinline__ void save_build_atom(CBox& b, uint32_t i) {
    if (b.transpose && i <= b.st->nlocal + b.st->nghost) b.tcount[i] = 0u;
    // the build sort counts from zero unless it reuses the spatial sort's
    // local counts (k_sort_zero's work)
    if (!b.st->local_cells_valid && !(b.st->err & ERR_GHOST_CAPACITY)) {
        const uint32_t stride = ((b.cap_atoms + 1u + BLK - 1u) / BLK) * BLK;
        const uint32_t nc = geom_of(b).ncells;
        for (uint32_t c = i; c <= nc; c += stride) b.counts[c] = 0u;
    }
    if (i >= b.st->nlocal || b.st->lb_fresh) return;
    b.lbx[i] = b.x[i];
    b.lby[i] = b.y[i];
    b.lbz[i] = b.z[i];
}

__global__ void __launch_bounds__(BLK) k_save_build(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        if (!b.st->rebuild_now) continue;
        const uint32_t i = tl.start + threadIdx.x;
        save_build_atom(b, i);
    }
}

// The rebuild's remaining ghost fields (k_ghost_fill) or, when the step does
// not rebuild, the forward refresh (k_forward), with save_build_positions
// (k_save_build), in one launch over every atom slot: the three touch
// different fields, and none reads what another writes.
__global__ void __launch_bounds__(BLK) k_ghosts_save(KARGS) {
    FOR_TILES_SPREAD(t) {  // the ghosts are the last tiles of each box
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        const uint32_t a = tl.start + threadIdx.x;
        const uint32_t nl = st->nlocal;
        if (a < nl || a >= nl + st->nghost) continue;
        if (!st->rebuild_now) forward_ghost(b, st, a);
        else if (!(st->err & ERR_GHOST_CAPACITY)) fill_ghost(b, st, a);
    }
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        if (b.st->rebuild_now) save_build_atom(b, tl.start + threadIdx.x);
    }
}

// ── CSR build: neighbor.rs:1175-1304 ──────────────────────────────────────
// Row i = accepted candidates of the self cell after atom i (newton; the whole
// cell but i without newton), then of each stencil cell in stencil order, each
// cell in bin order. The tests are the reference's: r2 = fma(dx,dx,fma(dy,dy,dz*dz))
// against (2 r sf)^2 (uniform) or (ri+rj)^2 sf^2.

template <bool UNIFORM>
__device__ __forceinline__ bool accept_pair(double xi, double yi, double zi, double si, const double4& q, double ucut,
                                            double sfsq) {
    const double dx = q.x - xi;
    const double dy = q.y - yi;
    const double dz = q.z - zi;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    if (UNIFORM) return r2 < ucut;
    const double sum = si + q.w;
    return r2 < sum * sum * sfsq;
}

// Global-memory row walk, used by tiles whose neighborhood does not fit LDS.
template <bool FILL, bool UNIFORM>
__device__ uint32_t csr_row_g(CBox& b, CGeom& g, uint32_t i, GMEM uint32_t* out, uint32_t out_cap = 0xffffffffu) {
    const uint32_t c = b.cell[i];
    const double4 pi = ((const double4*)b.xr)[i];
    const double4* __restrict__ sp4 = (const double4*)b.spos;
    const uint32_t skip = g.newton ? 0xffffffffu : i;
    uint32_t n = 0;
    const uint32_t cap = out_cap;
    auto take = [&](uint32_t m) {
        const uint32_t j = b.sorted[m];
        if (j != skip) {
            if (FILL && n < cap) out[n] = j;
            ++n;
        }
    };
    auto scan = [&](uint32_t m0, uint32_t m1) {
        uint32_t m = m0;
        for (; m + 1 < m1; m += 2) {  // two candidate loads in flight
            const double4 q0 = sp4[m], q1 = sp4[m + 1];
            const bool a0 = accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, q0, g.ucut_sq, g.sf_sq);
            const bool a1 = accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, q1, g.ucut_sq, g.sf_sq);
            if (a0) take(m);
            if (a1) take(m + 1);
        }
        if (m < m1 && accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, sp4[m], g.ucut_sq, g.sf_sq)) take(m);
    };
    // self cell, merged with a first run that starts at the next cell (newton)
    int r0 = 0;
    if (g.has_self) {
        if (g.newton) {
            uint32_t cend = c + 1;
            if (g.nruns > 0 && g.run_off[0] == 1) {
                cend = c + 1 + (uint32_t)g.run_len[0];
                r0 = 1;
            }
            if (cend > g.ncells) cend = g.ncells;
            scan(b.sidx[i] + 1, b.bin_start[cend]);
        } else {
            scan(b.bin_start[c], b.bin_start[c + 1]);
        }
    }
    const int nr = g.nruns;
    const int32_t ncell = (int32_t)g.ncells;
    for (int k0 = r0; k0 < nr; k0 += 8) {
        uint32_t rs[8], re[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + q;
            rs[q] = 0u;
            re[q] = 0u;
            if (k < nr) {
                int32_t c0 = (int32_t)c + g.run_off[k];
                int32_t c1 = c0 + g.run_len[k];
                if (c0 < 0) c0 = 0;  // only atoms far outside the box reach past the grid
                if (c1 > ncell) c1 = ncell;
                if (c0 < c1) {
                    rs[q] = b.bin_start[c0];
                    re[q] = b.bin_start[c1];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) scan(rs[q], re[q]);
    }
    return n;
}

// LDS staging of one tile of BLK consecutive (cell-sorted) locals: the union
// of the cell ranges [c_lo + o, c_hi + o] over the stencil offsets o (and 0),
// merged in ascending order, copied from the bin-sorted arrays.
constexpr int CSR_LA = 2048;  // staged atoms
constexpr int CSR_LC = 1280;  // staged cell-table entries
constexpr int CSR_LR = 48;    // merged ranges
constexpr uint32_t CSR_OVERFLOW = 0xffffffffu;

struct CsrLds {
    double4 pos[CSR_LA];
    uint32_t j[CSR_LA];
    uint32_t ct[CSR_LC];
    int32_t rcs[CSR_LR], rce[CSR_LR];
    uint32_t rct[CSR_LR], rbase[CSR_LR], rlds[CSR_LR], rlen[CSR_LR];
    uint8_t kr[SOIL_MAX_STENCIL + 1];
    uint32_t nr, ok, c_lo, c_hi, unsorted;
    uint32_t scan[BLK / 64];
};

template <bool FILL, bool UNIFORM>
__device__ __forceinline__ uint32_t csr_row_lds(CBox& b, CGeom& g, const CsrLds& L, uint32_t i,
                                                uint32_t c, const double4& pi, uint32_t* out) {
    const uint32_t skip = g.newton ? 0xffffffffu : i;
    const int ns = g.nstencil;
    uint32_t n = 0;
    for (int e = g.has_self ? -1 : 0; e < ns; ++e) {
        const int kk = e < 0 ? ns : e;
        const uint32_t r = L.kr[kk];
        if (r == 255u) continue;
        const int32_t cc = (int32_t)c + (e < 0 ? 0 : g.stencil[e]);
        if (cc < L.rcs[r] || cc > L.rce[r]) continue;
        const uint32_t tix = L.rct[r] + (uint32_t)(cc - L.rcs[r]);
        uint32_t a0 = L.ct[tix];
        const uint32_t a1 = L.ct[tix + 1];
        if (e < 0 && g.newton) a0 = b.sidx[i] - L.rbase[r] + L.rlds[r] + 1u;
        for (uint32_t a = a0; a < a1; ++a) {
            if (accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, L.pos[a], g.ucut_sq, g.sf_sq)) {
                const uint32_t j = L.j[a];
                if (j != skip) {
                    if (FILL) out[n] = j;
                    ++n;
                }
            }
        }
    }
    return n;
}

// Pass 1: per tile, stage the neighborhood, count every row (offsets[i]) and
// emit the rows into the tile's slot of the temp region when they fit.
template <bool UNIFORM>
__device__ void csr_tile(CBox& b, CGeom& g, CsrLds& L, uint32_t i0) {
    const BoxState* st = b.st;
    const uint32_t nl = st->nlocal;
    const uint32_t tid = threadIdx.x;
    const uint32_t tix = i0 / BLK;
    const uint32_t ntl = (b.cap_local + BLK - 1) / BLK;
    const uint32_t tcap = b.cap_pairs / (ntl ? ntl : 1u);
    if (i0 >= nl) {  // only the CSR end marker
        if (tid == 0 && i0 == nl) {
            b.offsets[nl] = 0u;
            b.ctile[tix] = 0u;
        }
        return;
    }
    const uint32_t tn = nl - i0 < BLK ? nl - i0 : BLK;
    const uint32_t i = i0 + tid;
    const bool act = tid < tn;
    const uint32_t c = act ? b.cell[i] : 0u;
    if (tid == 0) {
        L.unsorted = 0u;
        L.ok = 1u;
        L.c_lo = c;
        L.c_hi = b.cell[i0 + tn - 1];
    }
    if (tn < BLK && tid == tn && i0 + tn == nl) b.offsets[nl] = 0u;
    __syncthreads();
    if (act && tid + 1 < tn && c > b.cell[i + 1]) L.unsorted = 1u;
    __syncthreads();
    if (tid == 0) {
        uint32_t ok = L.unsorted ? 0u : 1u;
        const int32_t ncell = (int32_t)g.ncells;
        const int32_t lo0 = (int32_t)L.c_lo, hi0 = (int32_t)L.c_hi;
        uint32_t nr = 0, ctot = 0;
        for (int q = 0; ok && q <= g.nstencil; ++q) {
            const int kk = g.sorder[q];
            const int32_t o = kk == g.nstencil ? 0 : g.stencil[kk];
            int32_t lo = lo0 + o, hi = hi0 + o;
            if (lo < 0) lo = 0;
            if (hi > ncell - 1) hi = ncell - 1;
            if (lo > hi) {
                L.kr[kk] = 255;
                continue;
            }
            if (nr > 0 && lo <= L.rce[nr - 1] + 1) {
                if (hi > L.rce[nr - 1]) L.rce[nr - 1] = hi;
            } else {
                if (nr == (uint32_t)CSR_LR) { ok = 0; break; }
                L.rcs[nr] = lo;
                L.rce[nr] = hi;
                ++nr;
            }
            L.kr[kk] = (uint8_t)(nr - 1);
        }
        for (uint32_t r = 0; ok && r < nr; ++r) {
            L.rct[r] = ctot;
            ctot += (uint32_t)(L.rce[r] - L.rcs[r]) + 2u;
        }
        if (ctot > (uint32_t)CSR_LC) ok = 0;
        L.nr = ok ? nr : 0u;
        L.ok = ok;
    }
    __syncthreads();
    if (tid < L.nr) {
        const uint32_t s0 = b.bin_start[L.rcs[tid]];
        L.rbase[tid] = s0;
        L.rlen[tid] = b.bin_start[L.rce[tid] + 1] - s0;
    }
    __syncthreads();
    if (tid == 0 && L.ok) {
        uint32_t tot = 0;
        for (uint32_t r = 0; r < L.nr; ++r) {
            L.rlds[r] = tot;
            tot += L.rlen[r];
        }
        if (tot > (uint32_t)CSR_LA) L.ok = 0;
    }
    __syncthreads();
    const double4 pi = act ? ((const double4*)b.xr)[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (!L.ok) {  // neighborhood too large for LDS: rows from global memory, filled in pass 2
        if (act) b.offsets[i] = csr_row_g<false, UNIFORM>(b, g, i, nullptr);
        if (tid == 0) b.ctile[tix] = CSR_OVERFLOW;
        __syncthreads();
        return;
    }
    const double4* __restrict__ sp4 = (const double4*)b.spos;
    for (uint32_t r = 0; r < L.nr; ++r) {
        const uint32_t base = L.rbase[r], dst = L.rlds[r], len = L.rlen[r];
        for (uint32_t x = tid; x < len; x += BLK) {
            L.pos[dst + x] = sp4[base + x];
            L.j[dst + x] = b.sorted[base + x];
        }
        const uint32_t ncr = (uint32_t)(L.rce[r] - L.rcs[r]) + 2u;
        for (uint32_t x = tid; x < ncr; x += BLK) L.ct[L.rct[r] + x] = b.bin_start[L.rcs[r] + x] - base + dst;
    }
    __syncthreads();
    const uint32_t cnt = act ? csr_row_lds<false, UNIFORM>(b, g, L, i, c, pi, nullptr) : 0u;
    uint32_t tot;
    const uint32_t lo = block_excl_scan(cnt, L.scan, tot);
    if (act) b.offsets[i] = cnt;
    if (tot <= tcap) {
        if (act && cnt) csr_row_lds<true, UNIFORM>(b, g, L, i, c, pi, b.ttmp + (size_t)tix * tcap + lo);
        if (tid == 0) b.ctile[tix] = tot;
    } else if (tid == 0) {
        b.ctile[tix] = CSR_OVERFLOW;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(BLK) k_csr_count(KARGS) {
    __shared__ CsrLds L;
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        CGeom& g = geom_of(b);
        if (g.uniform) csr_tile<true>(b, g, L, tl.start);
        else csr_tile<false>(b, g, L, tl.start);
    }
}

// Global-walk count (no LDS): mode 1 counts only; mode 2 also keeps rows of up
// to `slot` entries in a per-row slot of the temp region for the fill pass.
__global__ void __launch_bounds__(BLK) k_csr_count_g(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        CGeom& g = geom_of(b);
        const uint32_t i = tl.start + threadIdx.x;
        const uint32_t nl = st->nlocal;
        if (i > nl) continue;
        if (i == nl) {
            b.offsets[nl] = 0u;
            continue;
        }
        const uint32_t slot = b.csr_mode == 2 ? b.cap_pairs / (b.cap_local ? b.cap_local : 1u) : 0u;
        uint32_t n;
        if (slot) {
            GMEM uint32_t* out = b.ttmp + (size_t)i * slot;
            n = g.uniform ? csr_row_g<true, true>(b, g, i, out, slot) : csr_row_g<true, false>(b, g, i, out, slot);
        } else {
            n = g.uniform ? csr_row_g<false, true>(b, g, i, nullptr) : csr_row_g<false, false>(b, g, i, nullptr);
        }
        b.offsets[i] = n;
    }
}

// Wave-staged count (mode 3). A wave's 64 consecutive locals lie in
// nondecreasing cells (the step's spatial sort), so for each stencil run the
// union of their candidate ranges is one contiguous span of the bin-sorted
// records. The wave stages that span through LDS in coalesced chunks and each
// lane tests its own sub-range there; rows come out exactly as csr_row_g
// writes them (self cell, then the runs in order, each range ascending), into
// the same per-row slots. A wave whose cells are not sorted, or a run whose
// union is very wide, walks global memory per lane instead.
constexpr uint32_t CSRW_CH = 128;         // staged records per chunk and wave
constexpr uint32_t CSRW_MAX_SPAN = 1024;  // wider unions are walked per lane
static_assert(CSRW_CH % 64u == 0u, "a chunk is staged 64 records (one per lane) at a time");

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS (the tile-staged build, mode 5): the row's first CSR1_SLOTS accepted
// candidates are kept as record indices (positions in the bin-sorted order),
// slot k of lane l at lrow[k * 64 + l], and one word after the slots that
// takes the wave's discarded writes; the count goes on past them. Otherwise
// the row's first `slot` neighbours go to `out` (global row slots).
constexpr uint32_t CSR1_SLOTS = 10;
// the tile-staged count reuses a wave's slots as the neighbour-index staging of a walk
static_assert(CSR1_SLOTS * 64u >= CSRW_CH, "the slots must hold a staged chunk");

template <bool UNIFORM, bool NEWTON, bool LDS = false>
__device__ uint32_t csr_row_w(CBox& b, CGeom& g, uint32_t i, bool act, uint32_t c, uint32_t c_first,
                              uint32_t c_last, bool staged, uint32_t lane, double4* __restrict__ lpos,
                              uint32_t* __restrict__ lj, GMEM uint32_t* out, uint32_t slot,
                              uint32_t* __restrict__ lrow = nullptr) {
    const double4* __restrict__ sp4 = (const double4*)b.spos;
    const double4 pi = act ? ((const double4*)b.xr)[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int32_t ncell = (int32_t)g.ncells;
    // the full list skips the row atom itself: its own record
    const uint32_t self_m = (NEWTON || !act) ? 0xffffffffu : b.sidx[i];
    uint32_t n = 0;
    auto take = [&](uint32_t j) {
        if (NEWTON || j != i) {  // the full list skips the row atom itself
            if (n < slot) out[n] = j;
            ++n;
        }
    };
    // branch-free: every candidate is written to the lane's next slot (the
    // discard word once the slots are full) and kept by moving the slot
    // pointer on; a rejected one is overwritten by the next candidate. The
    // row's count is the pointer's distance at the end.
    uint32_t* sp = LDS ? lrow + lane : nullptr;
    uint32_t* const spmax = LDS ? lrow + CSR1_SLOTS * 64u : nullptr;
    auto take_m = [&](bool acc, uint32_t m) {
        if (!NEWTON) acc = acc && m != self_m;
        *(sp < spmax ? sp : spmax) = m;
        sp += acc ? 64u : 0u;
    };
    // this lane's range [r0, r1) inside the wave's union [u0, u1) (wave-uniform)
    auto span = [&](uint32_t r0, uint32_t r1, uint32_t u0, uint32_t u1) {
        if (!staged || u1 - u0 > CSRW_MAX_SPAN) {
            for (uint32_t m = r0; m < r1; ++m) {
                const bool acc = accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, sp4[m], g.ucut_sq, g.sf_sq);
                if (LDS) take_m(acc, m);
                else if (acc) take(b.sorted[m]);
            }
            return;
        }
        for (uint32_t base = u0; base < u1; base += CSRW_CH) {  // empty when u1 <= u0
            // all of the chunk's loads in flight at once: lanes past the
            // union's end load its last record again (staging never read)
            double4 v[CSRW_CH / 64u];
            uint32_t jv[CSRW_CH / 64u];
#pragma unroll
            for (uint32_t q = 0; q < CSRW_CH; q += 64u) {
                const uint32_t m = base + q + lane < u1 ? base + q + lane : u1 - 1u;
                v[q / 64u] = sp4[m];
                if (!LDS) jv[q / 64u] = b.sorted[m];
            }
#pragma unroll
            for (uint32_t q = 0; q < CSRW_CH; q += 64u) {
                lpos[q + lane] = v[q / 64u];
                if (!LDS) lj[q + lane] = jv[q / 64u];
            }
            wave_sync();
            const uint32_t lo = r0 > base ? r0 : base;
            const uint32_t hi = r1 < base + CSRW_CH ? r1 : base + CSRW_CH;
            if (LDS) {  // two candidates per pass: both LDS reads in flight before the tests
                uint32_t m = lo;
                const double4* lp = lpos + (lo - base);
                uint32_t m1 = m + 1u;
                while (m1 < hi) {
                    const double4 q0 = lp[0], q1 = lp[1];
                    const bool a0 = accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, q0, g.ucut_sq, g.sf_sq);
                    const bool a1 = accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, q1, g.ucut_sq, g.sf_sq);
                    take_m(a0, m);
                    take_m(a1, m1);
                    lp += 2;
                    m += 2u;
                    m1 = m + 1u;
                }
                if (m < hi) take_m(accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, lp[0], g.ucut_sq, g.sf_sq), m);
            } else {
                for (uint32_t m = lo; m < hi; ++m)
                    if (accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, lpos[m - base], g.ucut_sq, g.sf_sq))
                        take(lj[m - base]);
            }
            wave_sync();  // the next chunk overwrites the staging
        }
    };
    auto clampc = [&](int32_t x) { return x < 0 ? 0 : (x > ncell ? ncell : x); };
    int r0 = 0;
    if (g.has_self) {
        if (NEWTON) {
            int32_t extra = 1;
            if (g.nruns > 0 && g.run_off[0] == 1) {
                extra = 1 + g.run_len[0];
                r0 = 1;
            }
            const uint32_t ce = (uint32_t)clampc((int32_t)c + extra);
            const uint32_t ue = (uint32_t)clampc((int32_t)c_last + extra);
            span(act ? b.sidx[i] + 1u : 0u, act ? b.bin_start[ce] : 0u, b.bin_start[c_first], b.bin_start[ue]);
        } else {
            span(act ? b.bin_start[c] : 0u, act ? b.bin_start[c + 1] : 0u, b.bin_start[c_first],
                 b.bin_start[c_last + 1]);
        }
    }
    for (int k = r0; k < g.nruns; ++k) {
        const int32_t off = g.run_off[k], len = g.run_len[k];
        const int32_t uc0 = clampc((int32_t)c_first + off), uc1 = clampc((int32_t)c_last + off + len);
        if (staged && uc0 >= uc1) continue;  // wave-uniform: no lane has candidates
        int32_t c0 = (int32_t)c + off, c1 = c0 + len;
        if (c0 < 0) c0 = 0;
        if (c1 > ncell) c1 = ncell;
        uint32_t rs = 0u, re = 0u;
        if (act && c0 < c1) {
            rs = b.bin_start[c0];
            re = b.bin_start[c1];
        }
        span(rs, re, b.bin_start[uc0], b.bin_start[uc1]);
    }
    if (LDS) n = (uint32_t)(sp - (lrow + lane)) / 64u;
    return n;
}

__global__ void __launch_bounds__(BLK) k_csr_count_w(KARGS) {
    __shared__ double4 lpos[BLK / 64][CSRW_CH];
    __shared__ uint32_t lj[BLK / 64][CSRW_CH];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        CGeom& g = geom_of(b);
        const uint32_t nl = st->nlocal;
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        const uint32_t i0w = tl.start + wid * 64u;
        const uint32_t i = i0w + lane;
        if (i == nl) b.offsets[nl] = 0u;
        if (i0w >= nl) continue;  // wave-uniform; no block barrier in this kernel
        const bool act = i < nl;
        const uint32_t c = act ? b.cell[i] : 0u;
        const uint32_t last = (nl - i0w < 64u ? nl - i0w : 64u) - 1u;
        const uint32_t c_first = __shfl(c, 0, 64);
        const uint32_t c_last = __shfl(c, (int)last, 64);
        const uint32_t c_next = __shfl_down(c, 1, 64);
        const bool sorted = __ballot(act && lane < last && c > c_next) == 0ull;
        const uint32_t slot = b.cap_pairs / (b.cap_local ? b.cap_local : 1u);
        GMEM uint32_t* out = b.ttmp + (size_t)(act ? i : 0u) * slot;
        uint32_t n;
        if (g.newton)
            n = g.uniform ? csr_row_w<true, true>(b, g, i, act, c, c_first, c_last, sorted, lane, lpos[wid], lj[wid], out, slot)
                          : csr_row_w<false, true>(b, g, i, act, c, c_first, c_last, sorted, lane, lpos[wid], lj[wid], out, slot);
        else
            n = g.uniform ? csr_row_w<true, false>(b, g, i, act, c, c_first, c_last, sorted, lane, lpos[wid], lj[wid], out, slot)
                          : csr_row_w<false, false>(b, g, i, act, c, c_first, c_last, sorted, lane, lpos[wid], lj[wid], out, slot);
        if (act) b.offsets[i] = n;
    }
}

// ── CSR build, mode 4: flattened per-wave candidates ──────────────────────
// A wave takes 64 consecutive local rows. A row's candidates are a few
// contiguous ranges (segments) of the bin-sorted records: the self cell after
// the row's atom (under newton merged with a first run at offset +1; the whole
// cell without newton), then each stencil run, clamped to the grid. Laid end
// to end, row after row and segment after segment, they form one flattened
// index space per wave, in the order a row's neighbours are appended (self
// cell, then the stencil cells in order, each in bin order). Lane l tests the
// flattened candidates l, l+64, ...: every lane stays busy whatever the rows'
// lengths. A lane finds its segment by one mbcnt over a bitmap of segment
// starts and reads the segment's row and first record from LDS. Accepted
// candidates are compacted by ballot into the wave's dense region of the pair
// arena, which then holds the wave's rows in CSR order; per-row counts are
// LDS counters, and each row's offset within its wave goes to `offsets`. The
// offset scan runs over the waves' totals (wcount) and the fill copies each
// region to its place.
constexpr uint32_t FLAT_WORDS = 64;  // segment-start bitmap words per wave: 4096 flattened candidates
constexpr uint32_t FLAT_SEGS = 384;  // non-empty segments per wave

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v, uint32_t lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += y;
    }
    return v;
}

// Segment k of row i in cell c: k = -1 the self cell, k >= r0 the stencil runs.
template <bool NEWTON>
__device__ __forceinline__ void flat_segment(CBox& b, CGeom& g, uint32_t i, uint32_t c, int k, int32_t self_extra,
                                             uint32_t& m0, uint32_t& m1) {
    const int32_t ncell = (int32_t)g.ncells;
    if (k < 0) {
        if (NEWTON) {
            int32_t ce = (int32_t)c + self_extra;
            if (ce > ncell) ce = ncell;
            m0 = b.sidx[i] + 1u;
            m1 = b.bin_start[ce];
        } else {
            m0 = b.bin_start[c];
            m1 = b.bin_start[c + 1];
        }
        if (m1 < m0) m1 = m0;
        return;
    }
    int32_t c0 = (int32_t)c + g.run_off[k], c1 = c0 + g.run_len[k];
    if (c0 < 0) c0 = 0;  // only atoms far outside the box reach past the grid
    if (c1 > ncell) c1 = ncell;
    if (c0 < c1) {
        m0 = b.bin_start[c0];
        m1 = b.bin_start[c1];
    } else {
        m0 = m1 = 0u;
    }
}

// One row walked by its own lane over its segments (the wave's tables cannot
// hold it): the row's entries in the same order, into out when FILL.
template <bool FILL, bool UNIFORM, bool NEWTON>
__device__ uint32_t csr_row_segments(CBox& b, CGeom& g, uint32_t i, uint32_t c, int kfirst, int r0,
                                     int32_t self_extra, GMEM uint32_t* out) {
    const double4* __restrict__ sp4 = (const double4*)b.spos;
    const double4 pi = ((const double4*)b.xr)[i];
    uint32_t n = 0;
    for (int k = kfirst; k < g.nruns; k = (k < 0 ? r0 : k + 1)) {
        uint32_t m0, m1;
        flat_segment<NEWTON>(b, g, i, c, k, self_extra, m0, m1);
        for (uint32_t m = m0; m < m1; ++m) {
            if (!accept_pair<UNIFORM>(pi.x, pi.y, pi.z, pi.w, sp4[m], g.ucut_sq, g.sf_sq)) continue;
            const uint32_t j = b.sorted[m];
            if (!NEWTON && j == i) continue;
            if (FILL) out[n] = j;
            ++n;
        }
    }
    return n;
}

template <bool UNIFORM, bool NEWTON>
__device__ void csr_wave_flat(CBox& b, CGeom& g, uint32_t nl, uint32_t i0w, uint32_t lane, uint64_t* __restrict__ lbits,
                              uint64_t* __restrict__ lseg, double4* __restrict__ lrow, uint32_t* __restrict__ lcnt) {
    const uint32_t i = i0w + lane;
    const bool act = i < nl;
    const uint32_t c = act ? b.cell[i] : 0u;
    const uint32_t w = i0w >> 6;
    const uint32_t slot = b.cap_pairs / (b.cap_local ? b.cap_local : 1u);
    const size_t rstart = (size_t)w * 64u * slot;
    const uint32_t wcap = (uint32_t)min((size_t)64u * slot, (size_t)b.cap_pairs - rstart);
    GMEM uint32_t* region = b.ttmp + rstart;
    // the first run merges into the self cell when it starts at the next cell (newton)
    int r0 = 0;
    int32_t self_extra = 1;
    if (NEWTON && g.nruns > 0 && g.run_off[0] == 1) {
        self_extra = 1 + g.run_len[0];
        r0 = 1;
    }
    const int kfirst = g.has_self ? -1 : r0;
    // this row's flattened length and non-empty segments
    uint32_t T = 0, q = 0;
    if (act) {
        for (int k = kfirst; k < g.nruns; k = (k < 0 ? r0 : k + 1)) {
            uint32_t m0, m1;
            flat_segment<NEWTON>(b, g, i, c, k, self_extra, m0, m1);
            T += m1 - m0;
            q += m1 > m0 ? 1u : 0u;
        }
    }
    const uint64_t incl = wave_incl_scan64(((uint64_t)q << 32) | T, lane);
    const uint64_t tot64 = __shfl(incl, 63, 64);
    const uint32_t Ttot = (uint32_t)tot64, Qtot = (uint32_t)(tot64 >> 32);
    uint32_t nout;
    if (Ttot > FLAT_WORDS * 64u || Qtot > FLAT_SEGS) {
        // too wide for the wave's tables: each lane walks its own row (count,
        // then write at its offset within the wave)
        const uint32_t n =
            act ? csr_row_segments<false, UNIFORM, NEWTON>(b, g, i, c, kfirst, r0, self_extra, nullptr) : 0u;
        const uint64_t p = wave_incl_scan64(n, lane);
        const uint32_t loc = (uint32_t)p - n;
        nout = (uint32_t)__shfl(p, 63, 64);
        if (act) {
            b.offsets[i] = loc;
            if (nout <= wcap && n)
                csr_row_segments<true, UNIFORM, NEWTON>(b, g, i, c, kfirst, r0, self_extra, region + loc);
        }
    } else {
        // segment-start bitmap and segment table
        const uint32_t nwords = (Ttot + 63u) >> 6;
        for (uint32_t x = lane; x < nwords; x += 64u) lbits[x] = 0ull;
        lrow[lane] = act ? ((const double4*)b.xr)[i] : make_double4(0.0, 0.0, 0.0, 0.0);
        lcnt[lane] = 0u;
        wave_sync();
        if (act) {
            uint32_t F = (uint32_t)incl - T;           // this row's first flattened index
            uint32_t qk = (uint32_t)(incl >> 32) - q;  // and its first segment
            uint32_t* bits32 = (uint32_t*)lbits;
            for (int k = kfirst; k < g.nruns; k = (k < 0 ? r0 : k + 1)) {
                uint32_t m0, m1;
                flat_segment<NEWTON>(b, g, i, c, k, self_extra, m0, m1);
                if (m1 > m0) {
                    atomicOr(&bits32[F >> 5], 1u << (F & 31u));
                    lseg[qk++] = ((uint64_t)lane << 32) | (uint32_t)(m0 - F);
                    F += m1 - m0;
                }
            }
        }
        wave_sync();
        const double4* __restrict__ sp4 = (const double4*)b.spos;
        nout = 0;
        uint32_t base_rank = 0;
        for (uint32_t f0 = 0; f0 < Ttot; f0 += 64u) {
            const uint64_t word = lbits[f0 >> 6];
            const uint32_t f = f0 + lane;
            const bool valid = f < Ttot;
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0u));
            const uint32_t rank = base_rank + below + (uint32_t)((word >> lane) & 1ull) - 1u;
            base_rank += (uint32_t)__popcll(word);
            bool acc = false;
            uint32_t row = 0, j = 0;
            if (valid) {
                const uint64_t sv = lseg[rank];
                row = (uint32_t)(sv >> 32);
                const uint32_t m = f + (uint32_t)sv;
                const double4 pr = lrow[row];
                acc = accept_pair<UNIFORM>(pr.x, pr.y, pr.z, pr.w, sp4[m], g.ucut_sq, g.sf_sq);
                if (acc) {
                    j = b.sorted[m];
                    if (!NEWTON && j == i0w + row) acc = false;  // the full list skips the row atom itself
                }
            }
            const uint64_t am = __ballot(acc);
            if (acc) {
                const uint32_t pos = nout + lane_prefix(am);
                if (pos < wcap) region[pos] = j;
                atomicAdd(&lcnt[row], 1u);
            }
            nout += (uint32_t)__popcll(am);
        }
        wave_sync();
        const uint32_t n = lcnt[lane];
        const uint64_t p = wave_incl_scan64(n, lane);
        if (act) b.offsets[i] = (uint32_t)p - n;
        wave_sync();  // the next tile rewrites the tables
    }
    if (lane == 0) b.wcount[w] = nout;
}

__global__ void __launch_bounds__(BLK) k_csr_count_flat(KARGS) {
    __shared__ uint64_t lbits[BLK / 64][FLAT_WORDS];
    __shared__ uint64_t lseg[BLK / 64][FLAT_SEGS];
    __shared__ double4 lrow[BLK / 64][64];
    __shared__ uint32_t lcnt[BLK / 64][64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        CGeom& g = geom_of(b);
        const uint32_t nl = st->nlocal;
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        const uint32_t i0w = tl.start + wid * 64u;
        // the waves' totals end with a zero entry (its scan is the pair count)
        if (i0w + lane == nl) b.wcount[(nl + 63u) >> 6] = 0u;
        if (i0w >= nl) continue;  // wave-uniform; no block barrier in this kernel
        if (g.newton) {
            if (g.uniform) csr_wave_flat<true, true>(b, g, nl, i0w, lane, lbits[wid], lseg[wid], lrow[wid], lcnt[wid]);
            else csr_wave_flat<false, true>(b, g, nl, i0w, lane, lbits[wid], lseg[wid], lrow[wid], lcnt[wid]);
        } else {
            if (g.uniform) csr_wave_flat<true, false>(b, g, nl, i0w, lane, lbits[wid], lseg[wid], lrow[wid], lcnt[wid]);
            else csr_wave_flat<false, false>(b, g, nl, i0w, lane, lbits[wid], lseg[wid], lrow[wid], lcnt[wid]);
        }
    }
}

// Mode 4's fill (after the scan of the waves' totals): each row's offset is
// its wave's base plus its offset within the wave; each wave's region is
// copied to its place, one entry per lane, taking the consumer's transpose
// slots. A wave whose rows overflowed its region is walked again, per lane.
__global__ void __launch_bounds__(BLK) k_csr_fill_flat(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY))) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        const uint32_t i0w = tl.start + wid * 64u;
        const uint32_t i = i0w + lane;
        if (i == nl) b.offsets[nl] = b.wcount[(nl + 63u) >> 6];
        if (i0w >= nl) continue;  // wave-uniform
        const bool act = i < nl;
        const uint32_t w = i0w >> 6;
        const uint32_t base = b.wcount[w];
        const uint32_t tot = b.wcount[w + 1] - base;
        const uint32_t slot = b.cap_pairs / (b.cap_local ? b.cap_local : 1u);
        const size_t rstart = (size_t)w * 64u * slot;
        const uint32_t wcap = (uint32_t)min((size_t)64u * slot, (size_t)b.cap_pairs - rstart);
        const uint32_t loc = act ? b.offsets[i] : 0u;
        if (act) b.offsets[i] = base + loc;
        const bool tr = transpose_active(b);
        if (tot <= wcap) {
            const GMEM uint32_t* region = b.ttmp + rstart;
            for (uint32_t e = lane; e < tot; e += 64u) {
                const uint32_t j = region[e];
                b.indices[base + e] = j;
                if (tr) b.tslot[base + e] = atomicAdd(&b.tcount[j], 1u);  // consumer transpose slot
            }
        } else if (act) {
            CGeom& g = geom_of(b);
            GMEM uint32_t* out = b.indices + base + loc;
            const uint32_t n = g.uniform ? csr_row_g<true, true>(b, g, i, out) : csr_row_g<true, false>(b, g, i, out);
            if (tr)
                for (uint32_t e = base + loc; e < base + loc + n; ++e) b.tslot[e] = atomicAdd(&b.tcount[b.indices[e]], 1u);
        }
    }
}

// Pass 2 (after the offset scan): move each tile's rows from the temp region
// to their place (one contiguous copy), or walk overflowed tiles' rows again.
// Per-row slots (modes 2 and 3): a wave's 64 rows own one contiguous output
// range, copied one entry per lane (each lane finds its entry's row among the
// wave's row starts in LDS); a row that overflowed its slots is walked again
// by its own lane.
__global__ void __launch_bounds__(BLK) k_csr_fill(KARGS) {
    __shared__ uint32_t lstart[BLK / 64][64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY))) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t i0 = tl.start;
        if (i0 >= nl) continue;
        if (b.csr_mode >= 2) {
            const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
            const uint32_t i0w = i0 + wid * 64u;
            if (i0w >= nl) continue;  // wave-uniform; no block barrier in this path
            const uint32_t i = i0w + lane;
            const bool act = i < nl;
            const uint32_t last = (nl - i0w < 64u ? nl - i0w : 64u) - 1u;
            const uint32_t slot = b.cap_pairs / (b.cap_local ? b.cap_local : 1u);
            const uint32_t off = act ? b.offsets[i] : 0u;
            const uint32_t end = act ? b.offsets[i + 1] : 0u;
            const bool tr = transpose_active(b);
            if (act && end - off > slot) {  // overflowed: walk the row again
                CGeom& g = geom_of(b);
                GMEM uint32_t* out = b.indices + off;
                if (g.uniform) csr_row_g<true, true>(b, g, i, out);
                else csr_row_g<true, false>(b, g, i, out);
                if (tr)
                    for (uint32_t e = off; e < end; ++e) b.tslot[e] = atomicAdd(&b.tcount[b.indices[e]], 1u);
            }
            const uint32_t o0 = __shfl(off, 0, 64), o1 = __shfl(end, (int)last, 64);
            uint32_t* ls = lstart[wid];
            ls[lane] = act ? off : o1;
            wave_sync();
            for (uint32_t e = o0 + lane; e < o1; e += 64u) {
                uint32_t r = 0;  // the last row starting at or before e holds it
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if (r + step <= last && ls[r + step] <= e) r += step;
                const uint32_t rs = ls[r];
                const uint32_t re = r < last ? ls[r + 1] : o1;
                if (re - rs > slot) continue;  // its own lane walks it
                const uint32_t j = b.ttmp[(size_t)(i0w + r) * slot + (e - rs)];
                b.indices[e] = j;
                if (tr) b.tslot[e] = atomicAdd(&b.tcount[j], 1u);  // consumer transpose slot
            }
            wave_sync();  // the next tile rewrites the row starts
            continue;
        }
        if (b.csr_mode != 0) {
            const uint32_t i = i0 + threadIdx.x;
            if (i >= nl) continue;
            CGeom& g = geom_of(b);
            const uint32_t off = b.offsets[i], n = b.offsets[i + 1] - off;
            const uint32_t slot = b.csr_mode >= 2 ? b.cap_pairs / (b.cap_local ? b.cap_local : 1u) : 0u;
            GMEM uint32_t* out = b.indices + off;
            const bool tr = transpose_active(b);
            if (n <= slot) {
                const GMEM uint32_t* src = b.ttmp + (size_t)i * slot;
                for (uint32_t k = 0; k < n; ++k) {
                    const uint32_t j = src[k];
                    out[k] = j;
                    if (tr) b.tslot[off + k] = atomicAdd(&b.tcount[j], 1u);  // consumer transpose slot
                }
            } else {
                if (g.uniform) csr_row_g<true, true>(b, g, i, out);
                else csr_row_g<true, false>(b, g, i, out);
                if (tr)
                    for (uint32_t k = 0; k < n; ++k) b.tslot[off + k] = atomicAdd(&b.tcount[out[k]], 1u);
            }
            continue;
        }
        const uint32_t tix = i0 / BLK;
        const uint32_t ct = b.ctile[tix];
        if (ct == CSR_OVERFLOW) {
            const uint32_t i = i0 + threadIdx.x;
            if (i >= nl) continue;
            CGeom& g = geom_of(b);
            GMEM uint32_t* out = b.indices + b.offsets[i];
            if (g.uniform) csr_row_g<true, true>(b, g, i, out);
            else csr_row_g<true, false>(b, g, i, out);
            continue;
        }
        const uint32_t ntl = (b.cap_local + BLK - 1) / BLK;
        const uint32_t tcap = b.cap_pairs / (ntl ? ntl : 1u);
        const uint32_t* src = b.ttmp + (size_t)tix * tcap;
        uint32_t* dst = b.indices + b.offsets[i0];
        for (uint32_t x = threadIdx.x; x < ct; x += BLK) dst[x] = src[x];
    }
}

// ── CSR build, mode 5 (default): tile-staged rows ─────────────────────────
// Count (k_csr_tcount): a tile's waves count their rows as k_csr_count_w
// does, keeping each row's first CSR1_SLOTS accepted candidates in LDS (as
// record indices); the block's exclusive scan gives each row's offset within
// the tile (offsets[i]) and the tile's total (ctile). The tile's rows then go
// densely, in CSR order, into the tile's own region of the temp arena: rows
// that fit their slots are copied one entry per lane, longer rows are walked
// again by their own lanes (the staging spans only their cells). A tile whose
// total exceeds its region stages nothing.
// Scan: an exclusive scan over the tile totals (a few thousand per 10^6 rows)
// gives each tile's base, and the build's pair count.
// Place (k_csr_tplace): each tile's region is copied to its base, one entry
// per thread, taking the consumer's transpose slots; rows' offsets become
// base + offset in tile. An unstaged tile's rows are walked again here.

// Region of tile `tix` in the temp arena: cap_pairs split evenly over the
// box's local tiles (at least row-slots entries per row).
__device__ __forceinline__ uint32_t tile_region(CBox& b) {
    const uint32_t ntl = b.cap_local / BLK + 1u;
    return b.cap_pairs / ntl;
}

template <bool UNIFORM, bool NEWTON>
__device__ __forceinline__ uint32_t csr1_count(CBox& b, CGeom& g, uint32_t i0w, uint32_t nl, uint32_t lane,
                                               double4* lpos, uint32_t* lrow) {
    const uint32_t i = i0w + lane;
    const bool act = i < nl;
    const uint32_t c = act ? b.cell[i] : 0u;
    const uint32_t last = (nl - i0w < 64u ? nl - i0w : 64u) - 1u;
    const uint32_t c_first = __shfl(c, 0, 64);
    const uint32_t c_last = __shfl(c, (int)last, 64);
    const uint32_t c_next = __shfl_down(c, 1, 64);
    const bool sorted = __ballot(act && lane < last && c > c_next) == 0ull;
    return csr_row_w<UNIFORM, NEWTON, true>(b, g, i, act, c, c_first, c_last, sorted, lane, lpos, nullptr, nullptr,
                                            0u, lrow);
}

// Walks again, by their own lanes, the wave's rows selected by `sel`, each
// written from out + its offset (loc); the staging covers only their cells.
template <bool UNIFORM, bool NEWTON>
__device__ __forceinline__ void csr1_rewalk(CBox& b, CGeom& g, uint32_t i0w, uint32_t nl, uint32_t lane, bool sel,
                                            GMEM uint32_t* out, uint32_t loc, double4* lpos, uint32_t* lj) {
    const uint32_t i = i0w + lane;
    const bool act = i < nl;
    const uint64_t sm = __ballot(sel);
    if (sm == 0ull) return;  // wave-uniform
    const uint32_t last = (nl - i0w < 64u ? nl - i0w : 64u) - 1u;
    const uint32_t c = act ? b.cell[i] : 0u;
    const uint32_t c_next = __shfl_down(c, 1, 64);
    const bool sorted = __ballot(act && lane < last && c > c_next) == 0ull;
    const uint32_t c_first = __shfl(c, (int)__builtin_ctzll(sm), 64);
    const uint32_t c_last = __shfl(c, 63 - (int)__builtin_clzll(sm), 64);
    (void)csr_row_w<UNIFORM, NEWTON>(b, g, i, sel, c, c_first, c_last, sorted, lane, lpos, lj, out + loc,
                                     sel ? 0xffffffffu : 0u);
}

template <bool UNIFORM, bool NEWTON>
__device__ __forceinline__ uint32_t csr1_tile_count(CBox& b, CGeom& g, uint32_t i0, uint32_t nl, uint32_t wid,
                                                    uint32_t lane, double4* lpos, uint32_t* lrow, uint32_t* lscan) {
    const uint32_t i0w = i0 + wid * 64u;
    const uint32_t i = i0w + lane;
    const bool act = i < nl;
    uint32_t n = 0;
    if (i0w < nl) n = csr1_count<UNIFORM, NEWTON>(b, g, i0w, nl, lane, lpos, lrow);  // wave-uniform
    uint32_t tot;
    const uint32_t loc = block_excl_scan(n, lscan, tot);
    if (act) b.offsets[i] = loc;
    const uint32_t tcap = tile_region(b);
    if (tot > tcap || i0w >= nl) return tot;  // block-uniform / wave-uniform: nothing staged
    GMEM uint32_t* region = b.ttmp + (size_t)(i0 / BLK) * tcap;
    // each lane writes its own row's slots, every neighbour-index load in
    // flight at once (a lane past its row's count loads record 0's)
    const uint32_t kn = act && n <= CSR1_SLOTS ? n : 0u;  // longer rows are walked again below
    uint32_t jv[CSR1_SLOTS];
#pragma unroll
    for (uint32_t k = 0; k < CSR1_SLOTS; ++k) jv[k] = b.sorted[k < kn ? lrow[k * 64u + lane] : 0u];
#pragma unroll
    for (uint32_t k = 0; k < CSR1_SLOTS; ++k)
        if (k < kn) region[loc + k] = jv[k];
    wave_sync();  // lrow becomes the walk's staging of neighbour indices
    csr1_rewalk<UNIFORM, NEWTON>(b, g, i0w, nl, lane, act && n > CSR1_SLOTS, region, loc, lpos, lrow);
    return tot;
}

__global__ void __launch_bounds__(BLK) k_csr_tcount(KARGS) {
    __shared__ double4 lpos[BLK / 64][CSRW_CH];
    __shared__ uint32_t lrow[BLK / 64][CSR1_SLOTS * 64 + 1];  // the last word discards
    __shared__ uint32_t lscan[BLK / 64];
    const uint32_t tid = threadIdx.x, wid = tid >> 6, lane = tid & 63u;
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        const BoxState* st = b.st;
        if (!st->rebuild_now || (st->err & ERR_GHOST_CAPACITY)) continue;
        const uint32_t nl = st->nlocal;
        const uint32_t i0 = tl.start;
        if (i0 > nl) continue;  // block-uniform
        CGeom& g = geom_of(b);
        uint32_t tot;
        if (g.newton) tot = g.uniform ? csr1_tile_count<true, true>(b, g, i0, nl, wid, lane, lpos[wid], lrow[wid], lscan)
                                      : csr1_tile_count<false, true>(b, g, i0, nl, wid, lane, lpos[wid], lrow[wid], lscan);
        else tot = g.uniform ? csr1_tile_count<true, false>(b, g, i0, nl, wid, lane, lpos[wid], lrow[wid], lscan)
                             : csr1_tile_count<false, false>(b, g, i0, nl, wid, lane, lpos[wid], lrow[wid], lscan);
        const uint32_t tix = i0 / BLK;
        if (tid == 0) {
            b.ctile[tix] = tot;
            if (tot) atomicAdd(&b.cgroup[tix / CSR1_GROUP], tot);  // zeroed by the build's cell scan
            // read by the pull transpose after the placement consumes local_cells_valid
            if (i0 == 0u) {
                b.st->sstart_build = st->local_cells_valid;
                b.st->tpull_now = tpull_on(b) ? 1u : 0u;
            }
        }
        __syncthreads();  // the next tile rewrites the LDS tables
    }
}

template <bool UNIFORM, bool NEWTON>
__device__ __forceinline__ void csr1_tile_rewalk(CBox& b, CGeom& g, uint32_t i0, uint32_t nl, uint32_t wid,
                                                 uint32_t lane, uint32_t base, const uint32_t* locs, double4* lpos,
                                                 uint32_t* lj) {
    const uint32_t i0w = i0 + wid * 64u;
    if (i0w >= nl) return;  // wave-uniform
    const uint32_t i = i0w + lane;
    csr1_rewalk<UNIFORM, NEWTON>(b, g, i0w, nl, lane, i < nl, b.indices + base, locs[lane], lpos, lj);
}

template <bool PUB>
__global__ void __launch_bounds__(BLK) k_csr_tplace(KARGS) {
    __shared__ double4 lpos[BLK / 64][CSRW_CH];
    __shared__ uint32_t lj[BLK / 64][CSRW_CH];
    __shared__ uint32_t lloc[BLK / 64][64];
    __shared__ uint64_t lsum[BLK / 64];
    const uint32_t tid = threadIdx.x, wid = tid >> 6, lane = tid & 63u;
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        const uint32_t nl = st->nlocal;
        const uint32_t i0 = tl.start;
        if (!st->rebuild_now || (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY)) || i0 > nl) {  // block-uniform
            if (PUB && i0 == 0u && tid == 0u)  // the box's state goes to the host mirror, built or not
                publish_vals(b, __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), st->npairs);
            continue;
        }
        // the tile's base: the totals of the earlier groups of CSR1_GROUP
        // tiles, then of the group's earlier tiles; the pair count: every group
        const uint32_t tix = i0 / BLK, grp = tix / CSR1_GROUP;
        const uint32_t ngrp = (nl / BLK) / CSR1_GROUP + 1u;  // the groups of tiles 0 .. nl / BLK
        uint32_t pre = 0, all = 0;
        for (uint32_t q = tid; q < ngrp; q += BLK) {
            const uint32_t v = b.cgroup[q];
            all += v;
            pre += q < grp ? v : 0u;
        }
        if (tid < CSR1_GROUP && grp * CSR1_GROUP + tid < tix) pre += b.ctile[grp * CSR1_GROUP + tid];
        const uint64_t sums = block_sum64(((uint64_t)all << 32) | pre, lsum);  // neither half exceeds 32 bits
        const uint32_t base = (uint32_t)sums, npairs = (uint32_t)(sums >> 32);
        if (i0 == 0u && tid == 0u) {  // the build's bookkeeping, once per box
            st->local_cells_valid = 0;  // consumed by this build (neighbor.rs:596)
            st->npairs = npairs;
            uint32_t err;  // the bits as they stand after this update
            if (npairs > b.cap_pairs) {
                err = atomicOr(&st->err, ERR_PAIR_CAPACITY) | ERR_PAIR_CAPACITY;
            } else {
                err = __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                CGeom& g = geom_of(b);
                const uint64_t na = (uint64_t)nl + st->nghost;
                const uint64_t per = 24u + (g.uniform ? 0u : 8u) + 4u;
                st->csr_bytes += per * na + 4u * ((uint64_t)g.ncells + 1u) + 4u * ((uint64_t)nl + 1u) +
                                 4u * (uint64_t)npairs;
                st->csr_builds += 1;
            }
            if (PUB) publish_vals(b, err, npairs);
        }
        if (i0 == 0u && tpull_build(b))  // the pull transpose's group sums start from zero (it runs next)
            for (uint32_t q = tid; q < b.cap_atoms / (BLK * TP_GROUP) + 2u; q += BLK) b.tgroup[q] = 0u;
        if (npairs > b.cap_pairs) continue;  // block-uniform: the host grows the pairs and builds again
        const uint32_t tot = b.ctile[tix];
        const uint32_t i = i0 + tid;
        uint32_t loc = 0;
        if (i < nl) {
            loc = b.offsets[i];
            b.offsets[i] = base + loc;
        } else if (i == nl) {
            b.offsets[nl] = base + tot;
        }
        const bool tr = tatomic_active(b);
        const uint32_t tcap = tile_region(b);
        GMEM uint32_t* dst = b.indices + base;
        if (tot <= tcap) {  // block-uniform: the staged region, one entry per thread
            const GMEM uint32_t* src = b.ttmp + (size_t)tix * tcap;
            uint32_t e = tid;
            // four entries per thread in flight (a dense tile has many passes)
            for (; e + 3u * BLK < tot; e += 4u * BLK) {
                uint32_t j[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) j[u] = src[e + u * BLK];
#pragma unroll
                for (int u = 0; u < 4; ++u) dst[e + u * BLK] = j[u];
                if (tr) {
                    uint32_t sl[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) sl[u] = atomicAdd(&b.tcount[j[u]], 1u);  // consumer transpose slots
#pragma unroll
                    for (int u = 0; u < 4; ++u) b.tslot[base + e + u * BLK] = sl[u];
                }
            }
            for (; e < tot; e += BLK) {
                const uint32_t j = src[e];
                dst[e] = j;
                if (tr) b.tslot[base + e] = atomicAdd(&b.tcount[j], 1u);  // consumer transpose slot
            }
            continue;
        }
        // unstaged: every row of the tile walked again, then its transpose slots
        lloc[wid][lane] = loc;
        CGeom& g = geom_of(b);
        if (g.newton) {
            if (g.uniform) csr1_tile_rewalk<true, true>(b, g, i0, nl, wid, lane, base, lloc[wid], lpos[wid], lj[wid]);
            else csr1_tile_rewalk<false, true>(b, g, i0, nl, wid, lane, base, lloc[wid], lpos[wid], lj[wid]);
        } else {
            if (g.uniform) csr1_tile_rewalk<true, false>(b, g, i0, nl, wid, lane, base, lloc[wid], lpos[wid], lj[wid]);
            else csr1_tile_rewalk<false, false>(b, g, i0, nl, wid, lane, base, lloc[wid], lpos[wid], lj[wid]);
        }
        if (!tr) continue;  // block-uniform
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the rows just written, read back by other threads
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        uint32_t e = tid;
        for (; e + 3u * BLK < tot; e += 4u * BLK) {
            uint32_t j[4], sl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) j[u] = dst[e + u * BLK];
#pragma unroll
            for (int u = 0; u < 4; ++u) sl[u] = atomicAdd(&b.tcount[j[u]], 1u);
#pragma unroll
            for (int u = 0; u < 4; ++u) b.tslot[base + e + u * BLK] = sl[u];
        }
        for (; e < tot; e += BLK) b.tslot[base + e] = atomicAdd(&b.tcount[dst[e]], 1u);
        __syncthreads();  // lloc is rewritten by the next tile
    }
}

// ── consumer transpose: incoming pairs of every atom ─────────────────────
// The harmonic consumer sums exact integers, so the order of a column's
// entries is free: arrival slots from atomics place them directly.
__global__ void __launch_bounds__(BLK) k_tcount(KARGS) {
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b) || !tatomic_active(b)) continue;
        const uint32_t i = tl.start + threadIdx.x;
        if (i >= b.st->nlocal) continue;
        for (uint32_t e = b.offsets[i]; e < b.offsets[i + 1]; ++e) b.tslot[e] = atomicAdd(&b.tcount[b.indices[e]], 1u);
    }
}

// One entry per lane over the wave's contiguous entries, as in k_csr_fill.
__global__ void __launch_bounds__(BLK) k_tplace(KARGS) {
    __shared__ uint32_t lstart[BLK / 64][64];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b) || !tatomic_active(b)) continue;
        const uint32_t nl = b.st->nlocal;
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        const uint32_t i0w = tl.start + wid * 64u;
        if (i0w >= nl) continue;  // wave-uniform
        const uint32_t i = i0w + lane;
        const bool act = i < nl;
        const uint32_t last = (nl - i0w < 64u ? nl - i0w : 64u) - 1u;
        const uint32_t off = act ? b.offsets[i] : 0u;
        const uint32_t o0 = __shfl(off, 0, 64), o1 = b.offsets[i0w + last + 1u];
        uint32_t* ls = lstart[wid];
        ls[lane] = act ? off : o1;
        wave_sync();
        auto owner = [&](uint32_t e) {
            uint32_t r = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1)
                if (r + step <= last && ls[r + step] <= e) r += step;
            return r;
        };
        // four entries per lane in flight: long rows (clustered boxes) give a
        // wave many passes, each a chain of dependent loads
        uint32_t e = o0 + lane;
        for (; e + 192u < o1; e += 256u) {
            uint32_t j[4], sl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                j[u] = b.indices[e + 64u * u];
                sl[u] = b.tslot[e + 64u * u];
            }
            uint32_t ts[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) ts[u] = b.tstart[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) b.tsrc[ts[u] + sl[u]] = i0w + owner(e + 64u * u);
        }
        for (; e < o1; e += 64u) b.tsrc[b.tstart[b.indices[e]] + b.tslot[e]] = i0w + owner(e);
        wave_sync();
    }
}

// ── consumer transpose, pull strategy (sparse boxes, default build) ───────
// Row i holds target j exactly when j's cell lies in one of i's windows: i's
// own cell or cells [ci + o, ci + o + len) of a stencil run (o, len). So the
// rows that can hold a target in cells [cl, ch] lie in cells [cl, ch] and
// [cl - o - len + 1, ch - o] for every run. The locals are in cell order, so
// each merged cell range is one range of rows, and walking the merged ranges
// sees every CSR entry with a target in [cl, ch] exactly once. A tile of 256
// targets walks its local targets' rows one row per thread and keeps the
// entries whose target is its own; a ghost target (ghosts are not in cell
// order) walks the rows of its single cell's ranges, which are disjoint since
// the runs and the own cell are disjoint sets of offsets. k_tpull_count
// counts each target's incoming pairs (LDS counters) and the tile's total
// (and per-group sums); k_tpull_place turns the totals into tile bases,
// writes tstart and places tsrc in each column in any order, as the consumer
// allows. No device-scope atomics but one per tile.

// first local row whose cell is >= c (locals are in cell order); the spatial
// sort's starts unless they are stale (a regrown arena), then a search
__device__ __forceinline__ uint32_t tp_row_start(CBox& b, uint32_t c, uint32_t nl, bool fast) {
    if (fast) return b.sstart[c];
    uint32_t lo = 0, hi = nl;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b.cell[mid] < c) lo = mid + 1u;
        else hi = mid;
    }
    return lo;
}

struct TpRanges {
    int32_t lo[TP_MAXR], hi[TP_MAXR];  // merged cell ranges, ascending
    uint32_t r0[TP_MAXR];              // first row of each
    uint32_t pre[TP_MAXR + 1];         // rows before each (pre[n] = all)
    uint32_t n;
};

// The merged row ranges of the tile's local targets [a0, a0 + nloc) (every
// thread of the block takes part; ends with the ranges visible to all).
__device__ void tp_setup(CBox& b, CGeom& g, uint32_t a0, uint32_t nloc, uint32_t nl, bool fast, TpRanges& R) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        uint32_t n = 0;
        if (nloc > 0) {
            const int32_t nc = (int32_t)g.ncells;
            const int32_t cl = (int32_t)b.cell[a0], ch = (int32_t)b.cell[a0 + nloc - 1u];
            auto add = [&](int32_t a, int32_t z) {
                if (a < 0) a = 0;
                if (z > nc - 1) z = nc - 1;
                if (a > z || n == (uint32_t)TP_MAXR) return;  // (set-up keeps pull boxes within TP_MAXR)
                uint32_t k = n++;  // insertion by start
                while (k > 0 && R.lo[k - 1] > a) {
                    R.lo[k] = R.lo[k - 1];
                    R.hi[k] = R.hi[k - 1];
                    --k;
                }
                R.lo[k] = a;
                R.hi[k] = z;
            };
            add(cl, ch);  // the targets' own cells
            for (int k = 0; k < g.nruns; ++k) add(cl - g.run_off[k] - g.run_len[k] + 1, ch - g.run_off[k]);
            uint32_t m = 0;  // merge overlapping and adjacent ranges
            for (uint32_t k = 0; k < n; ++k) {
                if (m > 0 && R.lo[k] <= R.hi[m - 1] + 1) {
                    if (R.hi[k] > R.hi[m - 1]) R.hi[m - 1] = R.hi[k];
                } else {
                    R.lo[m] = R.lo[k];
                    R.hi[m] = R.hi[k];
                    ++m;
                }
            }
            n = m;
        }
        R.n = n;
    }
    __syncthreads();
    const uint32_t n = R.n;
    if (tid < n) {
        const uint32_t r0 = tp_row_start(b, (uint32_t)R.lo[tid], nl, fast);
        R.r0[tid] = r0;
        R.pre[tid + 1] = tp_row_start(b, (uint32_t)R.hi[tid] + 1u, nl, fast) - r0;  // rows, summed below
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
        R.pre[0] = 0;
        for (uint32_t k = 1; k <= n; ++k) {
            s += R.pre[k];
            R.pre[k] = s;
        }
    }
    __syncthreads();
}

// Calls f(target - a0, row) for every CSR entry of the merged rows whose
// target is one of the tile's locals [a0, a0 + nloc); one row per thread.
template <typename F>
__device__ __forceinline__ void tp_walk_local(CBox& b, const TpRanges& R, uint32_t a0, uint32_t nloc, F&& f) {
    const uint32_t rows = R.pre[R.n];
    for (uint32_t x = threadIdx.x; x < rows; x += BLK) {
        uint32_t k = 0;
        while (R.pre[k + 1] <= x) ++k;
        const uint32_t r = R.r0[k] + (x - R.pre[k]);
        const uint32_t o0 = b.offsets[r], o1 = b.offsets[r + 1];
        for (uint32_t e = o0; e < o1; e += 4u) {  // up to four entries' loads in flight
            uint32_t j[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) j[u] = e + u < o1 ? b.indices[e + u] : 0xffffffffu;
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u)
                if (j[u] - a0 < nloc) f(j[u] - a0, r);
        }
    }
}

// Calls f(row) for every CSR entry whose target is the ghost a (in cell cg):
// the rows of its own cell, then of each run's backward cells.
template <typename F>
__device__ __forceinline__ void tp_walk_ghost(CBox& b, CGeom& g, uint32_t a, uint32_t nl, bool fast, F&& f) {
    const int32_t nc = (int32_t)g.ncells;
    const int32_t cg = (int32_t)b.cell[a];
    for (int k = -1; k < g.nruns; ++k) {
        int32_t lo = k < 0 ? cg : cg - g.run_off[k] - g.run_len[k] + 1;
        int32_t hi = k < 0 ? cg : cg - g.run_off[k];
        if (lo < 0) lo = 0;
        if (hi > nc - 1) hi = nc - 1;
        if (lo > hi) continue;
        const uint32_t r0 = tp_row_start(b, (uint32_t)lo, nl, fast);
        const uint32_t r1 = tp_row_start(b, (uint32_t)hi + 1u, nl, fast);
        for (uint32_t r = r0; r < r1; ++r) {
            const uint32_t o0 = b.offsets[r], o1 = b.offsets[r + 1];
            for (uint32_t e = o0; e < o1; ++e)
                if (b.indices[e] == a) f(r);
        }
    }
}

__global__ void __launch_bounds__(BLK) k_tpull_count(KARGS) {
    __shared__ TpRanges R;
    __shared__ uint32_t cnt[BLK];
    __shared__ uint32_t lsum[BLK / 64];
    const uint32_t tid = threadIdx.x;
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b) || !tpull_active(b)) continue;
        const BoxState* st = b.st;
        const uint32_t nl = st->nlocal, na = nl + st->nghost, a0 = tl.start;
        if (a0 >= na) continue;  // block-uniform
        CGeom& g = geom_of(b);
        const uint32_t an = na - a0 < BLK ? na - a0 : BLK;
        const uint32_t nloc = a0 < nl ? (nl - a0 < an ? nl - a0 : an) : 0u;
        const bool fast = st->sstart_build != 0u;
        cnt[tid] = 0u;
        tp_setup(b, g, a0, nloc, nl, fast, R);  // (its barriers also publish cnt's zeros)
        tp_walk_local(b, R, a0, nloc, [&](uint32_t k, uint32_t) { atomicAdd(&cnt[k], 1u); });
        const uint32_t a = a0 + tid;
        if (tid < an && a >= nl) {
            uint32_t c = 0;
            tp_walk_ghost(b, g, a, nl, fast, [&](uint32_t) { ++c; });
            cnt[tid] = c;
        }
        __syncthreads();
        const uint32_t c = tid < an ? cnt[tid] : 0u;
        if (tid < an) b.tcount[a] = c;
        uint32_t tot;
        (void)block_excl_scan(c, lsum, tot);
        if (tid == 0) {
            b.ttile[a0 / BLK] = tot;
            if (tot) atomicAdd(&b.tgroup[(a0 / BLK) / TP_GROUP], tot);  // zeroed by the CSR placement
        }
        __syncthreads();  // the next tile rewrites cnt and the ranges
    }
}

__global__ void __launch_bounds__(BLK) k_tpull_place(KARGS) {
    __shared__ TpRanges R;
    __shared__ uint32_t ts[BLK], cur[BLK];
    __shared__ uint32_t lsum[BLK / 64];
    __shared__ uint64_t lsum64[BLK / 64];
    const uint32_t tid = threadIdx.x;
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b) || !tpull_active(b)) continue;
        const BoxState* st = b.st;
        const uint32_t nl = st->nlocal, na = nl + st->nghost, a0 = tl.start;
        if (a0 > na) continue;  // block-uniform (the tile holding index na writes tstart[na])
        // the tile's base: the totals of the earlier groups of tiles, then of
        // the group's earlier tiles
        const uint32_t tix = a0 / BLK, grp = tix / TP_GROUP;
        uint32_t pre = 0;
        for (uint32_t q = tid; q < grp; q += BLK) pre += b.tgroup[q];
        if (tid < TP_GROUP && grp * TP_GROUP + tid < tix) pre += b.ttile[grp * TP_GROUP + tid];
        const uint32_t base = (uint32_t)block_sum64(pre, lsum64);  // below 2^32: the pair count
        const uint32_t an = na - a0 < BLK ? na - a0 : BLK;
        const uint32_t c = tid < an ? b.tcount[a0 + tid] : 0u;
        uint32_t tot;
        const uint32_t loc = block_excl_scan(c, lsum, tot) + base;
        if (tid <= an && a0 + tid <= na) b.tstart[a0 + tid] = loc;  // tstart[na] = every incoming pair
        if (an == 0u) continue;  // block-uniform
        CGeom& g = geom_of(b);
        const uint32_t nloc = a0 < nl ? (nl - a0 < an ? nl - a0 : an) : 0u;
        const bool fast = st->sstart_build != 0u;
        ts[tid] = loc;
        cur[tid] = 0u;
        tp_setup(b, g, a0, nloc, nl, fast, R);
        tp_walk_local(b, R, a0, nloc, [&](uint32_t k, uint32_t r) {
            const uint32_t slot = atomicAdd(&cur[k], 1u);
            b.tsrc[ts[k] + slot] = r;
        });
        const uint32_t a = a0 + tid;
        if (tid < an && a >= nl) {
            uint32_t k = loc;
            tp_walk_ghost(b, g, a, nl, fast, [&](uint32_t r) { b.tsrc[k++] = r; });
        }
        __syncthreads();  // the next tile rewrites the LDS tables
    }
}

// Harmonic soft-sphere test consumer (not reference code). For a CSR pair
// (i, j), F = k (ri+rj-d)/d * (xj - xi) when 0 < d < ri+rj, quantized to
// q = round_half_even(F * 2^32); i receives -q and j receives +q. Each atom
// sums its row and its incoming pairs in int64 (exact, so any order gives the
// same sum) and adds sum * 2^-32 (exact) to its f64 force. Both sides evaluate
// F with the row atom first, so they round identically.
__device__ __forceinline__ void pair_force(const double4& pi, const double4& pj, double k, double& fx, double& fy,
                                           double& fz) {
    const double dx = pj.x - pi.x;
    const double dy = pj.y - pi.y;
    const double dz = pj.z - pi.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rs = pi.w + pj.w;
    if (r2 < rs * rs && r2 > 0.0) {
        const double dist = sqrt(r2);
        const double s = k * (rs - dist) / dist;
        fx = s * dx;
        fy = s * dy;
        fz = s * dz;
    } else {
        fx = 0.0;
        fy = 0.0;
        fz = 0.0;
    }
}

constexpr double kFixScale = 4294967296.0;  // 2^32

// q = round_half_even(f * 2^32). FAST: the box's forces are below 2^19 in
// magnitude (host check: |k| (ri + rj) with margin), so |f 2^32| < 2^51 and
// adding 1.5 * 2^52 rounds it to the nearest integer, ties to even, which then
// sits in the low bits of the sum's representation: one add and one integer
// subtraction instead of the general conversion.
template <bool FAST>
__device__ __forceinline__ long long fix_q(double f) {
    if (FAST) {
        const double t = f * kFixScale + 6755399441055744.0;  // 1.5 * 2^52
        return (long long)__double_as_longlong(t) - 0x4338000000000000ll;
    }
    return __double2ll_rn(f * kFixScale);
}

// Adds each lane's (qx, qy, qz) into its row's LDS accumulator. A pass's
// lanes hold consecutive entries, so each row's lanes are contiguous; when a
// pass spans few rows (long rows: clustered boxes) the lanes of a row are
// summed first by a segmented scan and its last lane adds the sum, instead of
// 64 lanes contending for one accumulator. Integer sums: the same bits either
// way.
template <bool DENSE>
__device__ __forceinline__ void harmonic_add(unsigned long long (*acc)[3], uint32_t lane, bool valid, uint32_t r,
                                             long long qx, long long qy, long long qz) {
    const uint32_t key = valid ? r : 0xffffffffu;
    bool direct = !DENSE;
    if (DENSE) {
        const uint32_t kp = __shfl_up(key, 1, 64);
        direct = __popcll(__ballot(valid && (lane == 0u || kp != key))) > 8;  // many rows, little contention
    }
    if (direct) {  // wave-uniform
        if (valid) {
            atomicAdd(&acc[r][0], (unsigned long long)qx);
            atomicAdd(&acc[r][1], (unsigned long long)qy);
            atomicAdd(&acc[r][2], (unsigned long long)qz);
        }
        return;
    }
    // the lanes of the same row below each lane, as a mask per shift (one scan
    // of the keys serves the three components)
    uint64_t same = 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t ko = __shfl_up(key, o, 64);
        if (lane >= (uint32_t)o && ko == key) same |= 1ull << o;
    }
    const uint32_t kn = __shfl_down(key, 1, 64);
    const bool tail = valid && (lane == 63u || kn != key);
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        long long q = valid ? (c == 0 ? qx : (c == 1 ? qy : qz)) : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(q, o, 64);
            if (same & (1ull << o)) q += y;
        }
        if (tail) atomicAdd(&acc[r][c], (unsigned long long)q);
    }
}

// The consumer. Each atom's force is its row's pairs (as the row atom) and its
// incoming pairs (as the neighbour): a wave's 64 atoms own one contiguous
// range of row entries and one of incoming entries; each lane evaluates every
// 64th pair of them (its atom found by a binary search over the wave's
// starts in LDS) and adds the fixed-point force into the atom's LDS
// accumulator. The sums are integers, so any evaluation order gives the same
// bits, and the lanes stay busy whatever the rows' lengths.
template <bool FAST>
__device__ __forceinline__ void harmonic_wave(CBox& b, const BoxState* st, uint32_t a0, uint32_t lane, uint32_t last,
                                              uint32_t* ls, const double4* pw, unsigned long long (*acc)[3]) {
    const double k = b.harmonic_k;
    const GMEM double4* xr = (const GMEM double4*)b.xr;
    const uint32_t nl = st->nlocal, nall = nl + st->nghost;
    const uint32_t a = a0 + lane;
    auto owner = [&](uint32_t e) {
        uint32_t r = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1)
            if (r + step <= last && ls[r + step] <= e) r += step;
        return r;
    };
    // rows (locals): -q to the row atom
    if (a0 < nl) {
        const uint32_t rend = b.offsets[a0 + last + 1u < nl ? a0 + last + 1u : nl];
        ls[lane] = a < nl ? b.offsets[a] : rend;
        wave_sync();
        auto rows = [&](auto dense) {
            for (uint32_t e0 = ls[0]; e0 < rend; e0 += 64u) {  // wave-uniform
                const uint32_t e = e0 + lane;
                const bool v = e < rend;
                double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
                if (v) q = xr[b.indices[e]];
                const uint32_t r = v ? owner(e) : 0u;
                double fx, fy, fz;
                pair_force(pw[r], q, k, fx, fy, fz);
                harmonic_add<decltype(dense)::value>(acc, lane, v, r, -fix_q<FAST>(fx), -fix_q<FAST>(fy),
                                                     -fix_q<FAST>(fz));
            }
        };
        // rows longer than 8 entries on average: passes span few rows
        if (rend - ls[0] > 8u * (last + 1u)) rows(std::true_type{});
        else rows(std::false_type{});
        wave_sync();
    }
    // incoming pairs: +q to the neighbour, the row atom evaluated first
    const uint32_t tend = b.tstart[a0 + last + 1u];
    ls[lane] = a < nall ? b.tstart[a] : tend;
    wave_sync();
    auto incoming = [&](auto dense) {
        for (uint32_t e0 = ls[0]; e0 < tend; e0 += 64u) {  // wave-uniform
            const uint32_t e = e0 + lane;
            const bool v = e < tend;
            double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
            if (v) q = xr[b.tsrc[e]];
            const uint32_t r = v ? owner(e) : 0u;
            double fx, fy, fz;
            pair_force(q, pw[r], k, fx, fy, fz);
            harmonic_add<decltype(dense)::value>(acc, lane, v, r, fix_q<FAST>(fx), fix_q<FAST>(fy), fix_q<FAST>(fz));
        }
    };
    if (tend - ls[0] > 8u * (last + 1u)) incoming(std::true_type{});
    else incoming(std::false_type{});
    wave_sync();
}

// ASSIGN (fused step, forces were not zeroed): f = sum * 2^-32, which equals
// 0 + sum * 2^-32 bit for bit; otherwise f += sum * 2^-32.
template <bool ASSIGN>
__global__ void __launch_bounds__(BLK) k_harmonic(KARGS) {
    __shared__ uint32_t lstart[BLK / 64][64];
    __shared__ double4 lpos[BLK / 64][64];
    __shared__ unsigned long long lacc[BLK / 64][64][3];
    FOR_TILES(t) {
        const Tile tl = tiles[t];
        CBox& b = box_at(boxes, tl.box);
        if (!live(b) || !b.consumer) continue;
        const BoxState* st = b.st;
        const uint32_t nall = st->nlocal + st->nghost;
        const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        const uint32_t a0 = tl.start + wid * 64u;
        if (a0 >= nall) continue;  // wave-uniform; no block barrier in this kernel
        const uint32_t a = a0 + lane;
        const bool act = a < nall;
        if (st->rebuilds == 0 || (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY))) {
            if (ASSIGN && act) {
                b.fx[a] = 0.0;
                b.fy[a] = 0.0;
                b.fz[a] = 0.0;
            }
            continue;
        }
        const uint32_t last = (nall - a0 < 64u ? nall - a0 : 64u) - 1u;
        unsigned long long (*acc)[3] = lacc[wid];
        acc[lane][0] = acc[lane][1] = acc[lane][2] = 0ull;
        if (act) lpos[wid][lane] = ((const GMEM double4*)b.xr)[a];
        wave_sync();
        if (b.fix_fast) harmonic_wave<true>(b, st, a0, lane, last, lstart[wid], lpos[wid], acc);  // box-uniform
        else harmonic_wave<false>(b, st, a0, lane, last, lstart[wid], lpos[wid], acc);
        if (act) {
            const double inv = 1.0 / kFixScale;
            const double gx = (double)(long long)acc[lane][0] * inv;
            const double gy = (double)(long long)acc[lane][1] * inv;
            const double gz = (double)(long long)acc[lane][2] * inv;
            if (ASSIGN) {
                b.fx[a] = gx;
                b.fy[a] = gy;
                b.fz[a] = gz;
            } else {
                b.fx[a] = b.fx[a] + gx;
                b.fy[a] = b.fy[a] + gy;
                b.fz[a] = b.fz[a] + gz;
            }
        }
        wave_sync();  // the next tile rewrites the wave's LDS
    }
}

// The force hook's device view: this step's counts and the current buffers
// (the spatial sort swaps the permuted columns), one thread per box.
__global__ void k_sync_view(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!b.view) continue;
        soil_hip_device_view& v = *b.view;
        if (!live(b)) {  // a stopped box: the hook sees no atoms
            v.nlocal = v.nghost = v.npairs = v.rebuilt = 0;
            continue;
        }
        const BoxState* st = b.st;
        v.nlocal = st->nlocal;
        v.nghost = st->nghost;
        v.npairs = st->npairs;
        v.rebuilt = st->rebuild_now;
        v.cap_local = b.cap_local;
        v.cap_atoms = b.cap_atoms;
        v.cap_pairs = b.cap_pairs;
        v.ncols = (uint32_t)b.ncols;
        v.x = b.x; v.y = b.y; v.z = b.z;
        v.vx = b.vx; v.vy = b.vy; v.vz = b.vz;
        v.fx = b.fx; v.fy = b.fy; v.fz = b.fz;
        v.xr = b.xr;
        v.mass = b.mass; v.inv_mass = b.inv_mass; v.radius = b.rad;
        v.tag = b.tag; v.type = b.type; v.origin = b.origin;
        v.offsets = b.offsets;
        v.indices = b.indices;
        v.tstart = b.transpose ? b.tstart : nullptr;
        v.tsrc = b.transpose ? b.tsrc : nullptr;
        for (int c = 0; c < SOIL_HIP_MAX_COLS; ++c) {
            v.cols[c] = c < b.ncols ? b.cols[c].data : nullptr;
            v.col_width[c] = c < b.ncols ? b.cols[c].width : 0;
        }
        v.step = st->steps;
    }
}

// PEND (batches): the step's final_integration is left to the next step's
// kick-drift or the end of the call (k_final_pending). A box that ran out of
// ghost or pair capacity does not advance: the host grows it and redoes the
// step's tail. In a batch the box also freezes there (its step limit drops to
// its step count, so every later launch of the call skips it) until the host,
// which sees the error between runs, grows it and steps it up to the others.
template <bool PEND>
__global__ void k_step_end(KARGS) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        CBox& b = box_at(boxes, tiles[t].box);
        if (!live(b)) continue;
        BoxState* st = b.st;
        if (st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY)) {
            if (PEND) st->step_limit = st->steps;
            continue;
        }
        if (PEND) st->final_pending = 1u;
        st->steps += 1;
        st->ring_pos = (st->ring_pos + 1u) % b.ring_len;
    }
}

// Copies each box's error bits and list sizes to out[4*k..]: err, nghost, npairs, steps
__global__ void k_collect(KARGS, uint32_t* out) {
    FOR_TILES(t) {
        if (threadIdx.x != 0) continue;
        const uint32_t k = tiles[t].box;
        const BoxState* st = boxes[k]->st;
        out[4 * k] = st->err;
        out[4 * k + 1] = st->nghost;
        out[4 * k + 2] = st->npairs;
        out[4 * k + 3] = (uint32_t)st->steps;
    }
}

// Streams n doubles from src to dst: a copy of known size that calibrates the
// HBM traffic counters (FETCH_SIZE, WRITE_SIZE) against algorithmic bytes.
// one element per thread over a grid that covers the copy (on MI355X this
// plain form streams faster than a capped grid-stride loop)
__global__ void __launch_bounds__(BLK) k_copy_probe(const double* __restrict__ src, double* __restrict__ dst,
                                                    size_t n) {
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void copy_probe(const double* src, double* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, s, src, dst, n);
}

// ── launch helpers ─────────────────────────────────────────────────────────
template <typename... A>
static void launch(void (*k)(BoxDev* const*, const Tile*, uint32_t, A...), const TileSet& ts, BoxDev* const* boxes,
                   hipStream_t s, uint32_t maxgrid, A... args) {
    if (ts.n == 0) return;
    uint32_t grid = ts.n < maxgrid ? ts.n : maxgrid;
    if (grid >= 8) grid = (grid + 7u) & ~7u;
    hipLaunchKernelGGL(k, dim3(grid), dim3(BLK), 0, s, boxes, (const Tile*)ts.d, ts.n, args...);
}

template <int W, bool PUB = false>
static void run_scan(const Launch& L, const TileSet& scan_tiles) {
    launch(k_scan_reduce<W>, scan_tiles, L.boxes, L.stream, L.maxgrid);
    launch(k_scan_down<W, PUB>, scan_tiles, L.boxes, L.stream, L.maxgrid);
}

// ZERO: launch k_sort_zero (the fused step zeroes the counts in an earlier kernel)
template <int W, int SW, bool ZERO = true>
static void run_sort(const Launch& L, const TileSet& zero_tiles, const TileSet& item_tiles, const TileSet& scan_tiles) {
    if (ZERO) launch(k_sort_zero<W>, zero_tiles, L.boxes, L.stream, L.maxgrid);
    launch(k_sort_count<W>, item_tiles, L.boxes, L.stream, L.maxgrid);
    run_scan<SW>(L, scan_tiles);
    launch(k_sort_scatter<W>, item_tiles, L.boxes, L.stream, L.maxgrid);
    launch(k_sort_rank<W>, item_tiles, L.boxes, L.stream, L.maxgrid);
}

void phase_deform(const Launch& L) { launch(k_deform_init, L.t_local, L.boxes, L.stream, L.maxgrid); }
void phase_initial(const Launch& L) { launch(k_initial<false, 0>, L.t_local, L.boxes, L.stream, L.maxgrid); }
void phase_initial_decide(const Launch& L) {
    launch(k_initial<true, 0>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_decide<true>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_final_initial_decide(const Launch& L) {
    launch(k_initial<true, 1>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_decide<true>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_decide(const Launch& L) {
    launch(k_disp, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_decide<true>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
// the kick-drift (and the displacement test) alone, for steps whose decision
// waits on host work in between (shrink-wrap bounds)
void phase_kick(const Launch& L, bool with_final) {
    if (with_final) launch(k_initial<true, 1>, L.t_local, L.boxes, L.stream, L.maxgrid);
    else launch(k_initial<true, 0>, L.t_local, L.boxes, L.stream, L.maxgrid);
}
void phase_decide_fused(const Launch& L) { launch(k_decide<true>, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_extent(const Launch& L) {
    launch(k_extent_init, L.t_box, L.boxes, L.stream, L.maxgrid);
    launch(k_extent, L.t_local, L.boxes, L.stream, L.maxgrid);
}
void phase_decide_only(const Launch& L) {
    launch(k_disp, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_decide<false>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_remove_ghosts(const Launch& L) { launch(k_remove_ghosts, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_set_rebuild_now(const Launch& L) { launch(k_set_rebuild_now, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_zero(const Launch& L) { launch(k_zero<false>, L.t_atom, L.boxes, L.stream, L.maxgrid); }
void phase_zero_fused(const Launch& L) { launch(k_zero<true>, L.t_atom, L.boxes, L.stream, L.maxgrid); }
// pbc with the removals it implies (a box with no non-periodic axis removes none)
static void run_pbc(const Launch& L, bool open_only) {
    if (open_only) launch(k_pbc<true>, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else launch(k_pbc<false>, L.t_local1, L.boxes, L.stream, L.maxgrid);
    run_scan<SCAN_REMOVE>(L, L.t_scan_local);
    launch(k_remove_fill, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_remove_commit, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_pbc(const Launch& L) { run_pbc(L, false); }
void phase_sort(const Launch& L) {
    launch(k_pre_sort, L.t_box, L.boxes, L.stream, L.maxgrid);
    run_sort<SORT_SPATIAL, SCAN_CELLS_SPATIAL>(L, L.t_cell, L.t_local, L.t_scan_cell);
    launch(k_permute<false>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_swap_ptrs, L.t_box, L.boxes, L.stream, L.maxgrid);
    launch(k_post_sort<false>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
// sort + permute + (when rebuilding) the build's bookkeeping in one per-box kernel
void phase_pbc_sort_fused(const Launch& L) {
    if (L.any_open) run_pbc(L, true);  // boxes whose pbc may remove atoms
    launch(k_sort_count<SORT_SPATIAL, true>, L.t_local, L.boxes, L.stream, L.maxgrid);  // k_initial zeroed
    run_scan<SCAN_CELLS_SPATIAL>(L, L.t_scan_cell);
    launch(k_sort_scatter<SORT_SPATIAL>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_sort_rank<SORT_SPATIAL>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_permute<true>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_post_sort<true>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_borders(const Launch& L) {
    launch(k_border_count, L.t_border, L.boxes, L.stream, L.maxgrid);
    for (int d = 0; d < 3; ++d) launch(k_border_write, L.t_border, L.boxes, L.stream, L.maxgrid, d);
    launch(k_ghost_fill, L.t_ghost, L.boxes, L.stream, L.maxgrid);
}
void phase_forward(const Launch& L) { launch(k_forward, L.t_ghost, L.boxes, L.stream, L.maxgrid); }
void phase_build(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2) {
    launch(k_build_prep, L.t_box, L.boxes, L.stream, L.maxgrid);
    phase_build_fused(L, ev0, evc, ev1, ev2);
}
// the build after save_build: bins, CSR count, offsets, fill and transpose;
// with `publish` the offset scan (mode 5: the placement) writes the state
// into the host mirror
static void build_rest(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2, bool publish) {
    run_sort<SORT_BUILD, SCAN_CELLS_BUILD, false>(L, L.t_cell, L.t_atom, L.t_scan_cell);  // save_build zeroed
    if (ev0) (void)hipEventRecord(ev0, L.stream);
    if (L.csr_mode == 0) launch(k_csr_count, L.t_local1, L.boxes, L.stream, L.maxgrid);  // tiles + end marker
    else if (L.csr_mode == 3) launch(k_csr_count_w, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else if (L.csr_mode == 4) launch(k_csr_count_flat, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else if (L.csr_mode == 5) launch(k_csr_tcount, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else launch(k_csr_count_g, L.t_local1, L.boxes, L.stream, L.maxgrid);
    if (evc) (void)hipEventRecord(evc, L.stream);
    // the scan publishes the build's counts (over the waves' totals in mode 4,
    // the rows' counts in modes 0-3)
    if (L.csr_mode == 4) {
        if (publish) run_scan<SCAN_WAVES, true>(L, L.t_scan_local);
        else run_scan<SCAN_WAVES>(L, L.t_scan_local);
    } else if (L.csr_mode == 5) {
        // no scan: the placement sums the tiles' totals itself
    } else {
        if (publish) run_scan<SCAN_CSR, true>(L, L.t_scan_local);
        else run_scan<SCAN_CSR>(L, L.t_scan_local);
    }
    if (ev1) (void)hipEventRecord(ev1, L.stream);
    if (L.csr_mode == 4) launch(k_csr_fill_flat, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else if (L.csr_mode == 5 && publish) launch(k_csr_tplace<true>, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else if (L.csr_mode == 5) launch(k_csr_tplace<false>, L.t_local1, L.boxes, L.stream, L.maxgrid);
    else launch(k_csr_fill, L.t_local, L.boxes, L.stream, L.maxgrid);
    if (ev2) (void)hipEventRecord(ev2, L.stream);
    // consumer transpose (tcount was cleared by save_build; modes 1 and 2
    // take the arrival slots in k_csr_fill)
    if (L.tpull_any) {
        launch(k_tpull_count, L.t_atom, L.boxes, L.stream, L.maxgrid);
        launch(k_tpull_place, L.t_atom1, L.boxes, L.stream, L.maxgrid);
    }
    if (L.tatomic_any) {
        if (L.csr_mode == 0) launch(k_tcount, L.t_local, L.boxes, L.stream, L.maxgrid);
        run_scan<SCAN_TRANSPOSE>(L, L.t_scan_atom);
        launch(k_tplace, L.t_local, L.boxes, L.stream, L.maxgrid);
    }
}
// the build after phase_pbc_sort_fused (which already did the bookkeeping)
void phase_build_fused(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2) {
    launch(k_save_build, L.t_atom1, L.boxes, L.stream, L.maxgrid);
    build_rest(L, ev0, evc, ev1, ev2, false);
}
void phase_rebuild_fused(const Launch& L, hipEvent_t ev0, hipEvent_t evc, hipEvent_t ev1, hipEvent_t ev2, bool publish) {
    launch(k_border_count, L.t_border, L.boxes, L.stream, L.maxgrid);
    for (int d = 0; d < 3; ++d) launch(k_border_write, L.t_border, L.boxes, L.stream, L.maxgrid, d);
    launch(k_ghosts_save, L.t_atom1, L.boxes, L.stream, L.maxgrid);
    build_rest(L, ev0, evc, ev1, ev2, publish);
}
void phase_force(const Launch& L) { launch(k_harmonic<false>, L.t_atom, L.boxes, L.stream, L.maxgrid); }
void phase_force_fused(const Launch& L) { launch(k_harmonic<true>, L.t_atom, L.boxes, L.stream, L.maxgrid); }
void phase_reverse(const Launch& L) {
    // rounds last -> first, lo swap then hi swap (comm.rs:965-1021)
    for (int d = 2; d >= 0; --d) {
        launch(k_reverse, L.t_ghost, L.boxes, L.stream, L.maxgrid, d, 0);
        if (L.rev_split[d]) launch(k_reverse, L.t_ghost, L.boxes, L.stream, L.maxgrid, d, 1);
    }
}
void phase_final(const Launch& L) { launch(k_final, L.t_local, L.boxes, L.stream, L.maxgrid); }
void phase_ke(const Launch& L, const uint8_t* mask, double* part) {
    launch(k_ke, L.t_local, L.boxes, L.stream, L.maxgrid, mask, part);
}
void sum_tiles(const double* part, uint32_t n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(64), 0, s, part, n, out);
}

// ── fixed workspace of the allocation-free entry points ──────────────────
// One array per GPU, sized at compile time and held until the process ends.
// Host code never uses the name `scratch` as an address (on the host it is a
// stand-in symbol, not device memory): scratch_init looks up each device's
// copy once, with that device current, and the entry points pass that
// address to the kernels.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return -1;
    g_scratch_ptr[device] = p;
    return 0;
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}
void phase_classify(const Launch& L) { launch(k_classify, L.t_local, L.boxes, L.stream, L.maxgrid); }
void phase_zero_counts(const Launch& L) { launch(k_zero_counts, L.t_cell, L.boxes, L.stream, L.maxgrid); }
void phase_sync_view(const Launch& L) { launch(k_sync_view, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_step_end(const Launch& L) { launch(k_step_end<false>, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_initial_decide_pending(const Launch& L) {
    launch(k_initial<true, 2>, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_decide<true>, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_step_end_pending(const Launch& L) { launch(k_step_end<true>, L.t_box, L.boxes, L.stream, L.maxgrid); }
void phase_final_pending(const Launch& L) {
    launch(k_final_pending, L.t_local, L.boxes, L.stream, L.maxgrid);
    launch(k_clear_pending, L.t_box, L.boxes, L.stream, L.maxgrid);
}
void phase_collect(const Launch& L, uint32_t* out) { launch(k_collect, L.t_box, L.boxes, L.stream, L.maxgrid, out); }

}  // namespace soil_hip
