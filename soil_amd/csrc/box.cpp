// Box lifecycle, device arenas, the plain single-box step and the per-phase
// C ABI. One box = one HIP stream; all state lives on the device and the host
// reads it back lazily.

#include "box.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>

using namespace soil_hip;

namespace soil_hip {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIPCHK(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return set_error(SOIL_HIP_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// The harmonic consumer's pair force is at most |k| (ri + rj) per component;
// below 2^19 (with a 1 % margin) every q = F 2^32 is below 2^51 in magnitude
static int32_t fix_fast_ok(double k, double rmax) {
    return std::fabs(k) * 2.0 * rmax * 1.01 < 524288.0 ? 1 : 0;
}

static uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

static void push_tiles(std::vector<Tile>& v, uint32_t box, uint64_t n, uint32_t step) {
    for (uint64_t s = 0; s < n; s += step) v.push_back(Tile{box, (uint32_t)s});
}

void build_tiles(TileTables& tt, uint32_t b, const Caps& c) {
    tt.box.push_back(Tile{b, 0});
    push_tiles(tt.local, b, c.local, BLK);
    push_tiles(tt.local1, b, (uint64_t)c.local + 1, BLK);
    push_tiles(tt.atom, b, c.atoms, BLK);
    push_tiles(tt.atom1, b, (uint64_t)c.atoms + 1, BLK);
    push_tiles(tt.ghost, b, (uint64_t)c.atoms - c.local, BLK);
    push_tiles(tt.cell, b, (uint64_t)c.cells + 1, BLK);
    push_tiles(tt.border, b, c.atoms, BORDER_TILE);
    push_tiles(tt.scan_cell, b, (uint64_t)c.cells + 1, SCAN_TILE);
    push_tiles(tt.scan_local, b, (uint64_t)c.local + 1, SCAN_TILE);
    push_tiles(tt.scan_atom, b, (uint64_t)c.atoms + 1, SCAN_TILE);
}

// Tiles of one space are interleaved across boxes in XCD-friendly order: the
// tables keep box-major order; kernels stride over them with the grid.
int upload_tiles(const TileTables& tt, Launch& L, void** arena) {
    const std::vector<Tile>* lists[11] = {&tt.box,   &tt.local,  &tt.local1,    &tt.atom,       &tt.atom1,   &tt.ghost,
                                          &tt.cell,  &tt.border, &tt.scan_cell, &tt.scan_local, &tt.scan_atom};
    TileSet* sets[11] = {&L.t_box,  &L.t_local,  &L.t_local1,    &L.t_atom,       &L.t_atom1,   &L.t_ghost,
                         &L.t_cell, &L.t_border, &L.t_scan_cell, &L.t_scan_local, &L.t_scan_atom};
    size_t total = 0;
    for (auto* l : lists) total += l->size();
    if (*arena) (void)hipFree(*arena);
    *arena = nullptr;
    for (auto* s : sets) *s = TileSet{};
    if (total == 0) return 0;
    HIPCHK(hipMalloc(arena, total * sizeof(Tile)));
    Tile* base = (Tile*)*arena;
    size_t off = 0;
    for (int k = 0; k < 11; ++k) {
        if (!lists[k]->empty())
            HIPCHK(hipMemcpy(base + off, lists[k]->data(), lists[k]->size() * sizeof(Tile), hipMemcpyHostToDevice));
        sets[k]->d = base + off;
        sets[k]->n = (uint32_t)lists[k]->size();
        off += lists[k]->size();
    }
    return 0;
}

}  // namespace soil_hip

// ── arena layout ───────────────────────────────────────────────────────────
namespace soil_hip {

struct Carver {
    char* base;
    size_t off = 0;
    bool dry;
    template <typename T>
    void take(T*& p, size_t n) {
        off = (off + 255) & ~(size_t)255;
        p = dry ? nullptr : (T*)(base + off);
        off += std::max<size_t>(n, 1) * sizeof(T);
    }
};

void carve_atoms(Carver& c, BoxDev& d, const Caps& k, const std::vector<HostCol>& cols) {
    const size_t A = k.atoms, L1 = (size_t)k.local + 1, C1 = (size_t)k.cells + 1;
    c.take(d.x, A); c.take(d.y, A); c.take(d.z, A);
    c.take(d.vx, A); c.take(d.vy, A); c.take(d.vz, A);
    c.take(d.fx, A); c.take(d.fy, A); c.take(d.fz, A);
    c.take(d.mass, A); c.take(d.inv_mass, A); c.take(d.rad, A);
    c.take(d.tag, A); c.take(d.type, A); c.take(d.origin, A);
    c.take(d.imx, A); c.take(d.imy, A); c.take(d.imz, A);
    c.take(d.ax, A); c.take(d.ay, A); c.take(d.az, A);
    c.take(d.avx, A); c.take(d.avy, A); c.take(d.avz, A);
    c.take(d.amass, A); c.take(d.ainv_mass, A); c.take(d.arad, A);
    c.take(d.atag, A); c.take(d.atype, A);
    c.take(d.aimx, A); c.take(d.aimy, A); c.take(d.aimz, A);
    c.take(d.gswap, A);
    c.take(d.xr, 4 * A);
    c.take(d.lbx, k.local); c.take(d.lby, k.local); c.take(d.lbz, k.local);
    c.take(d.cell, A); c.take(d.slot, A); c.take(d.tmp, A); c.take(d.sorted, A); c.take(d.sidx, A);
    c.take(d.spos, 4 * A);
    c.take(d.bin_start, C1); c.take(d.counts, C1);
    c.take(d.scan_part, std::max(C1, A + 1) / SCAN_TILE + 2);
    c.take(d.red, (size_t)k.local / BLK + 2);
    c.take(d.scell, k.local); c.take(d.sstart, C1); c.take(d.perm, k.local);
    c.take(d.offsets, L1);
    c.take(d.tcount, A + 1); c.take(d.tstart, A + 1);
    c.take(d.ttile, A / BLK + 2); c.take(d.tgroup, A / (BLK * TP_GROUP) + 2);
    c.take(d.ctile, (size_t)k.local / BLK + 2);
    c.take(d.cgroup, (size_t)k.local / (BLK * CSR1_GROUP) + 2);
    c.take(d.wcount, (size_t)k.local / 64 + 2);
    c.take(d.bcnt, 6 * ((size_t)k.atoms / BORDER_TILE + 2));
    for (size_t i = 0; i < cols.size(); ++i) {
        c.take(d.cols[i].data, A * cols[i].width);
        c.take(d.cols[i].alt, A * cols[i].width);
    }
}

void carve_pairs(Carver& c, BoxDev& d, const Caps& k) {
    const size_t P = k.pairs;
    c.take(d.indices, P); c.take(d.ttmp, P); c.take(d.tslot, P); c.take(d.tsrc, P);
}

int alloc_arena(void** arena, size_t bytes) {
    if (*arena) (void)hipFree(*arena);
    *arena = nullptr;
    hipError_t e = hipMalloc(arena, bytes);
    if (e != hipSuccess) return set_error(SOIL_HIP_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return 0;
}

int push_desc(soil_hip_box* b) {
    HIPCHK(hipMemcpyAsync(b->d_dev, &b->h_dev, sizeof(BoxDev), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int pull_desc(soil_hip_box* b) {
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(&b->h_dev, b->d_dev, sizeof(BoxDev), hipMemcpyDeviceToHost));
    return 0;
}

int pull_state(soil_hip_box* b) {
    HIPCHK(hipMemcpyAsync(b->h_st, b->d_st, sizeof(BoxState), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->pub_expect = b->h_st->pub_seq;  // the stream is drained: the publish count is exact
    return 0;
}

// Waits for the publish the host enqueued last (pub_expect) and copies its
// fields (state, error bits, counts) into h_st. Falls back to a full state
// read when the stream drains without it (a stopped box publishes nothing)
// or the count is out of step.
int wait_mirror(soil_hip_box* b) {
    const uint64_t want = b->pub_expect;
    const HostMirror* m = b->h_mirror;
    for (uint32_t it = 1;; ++it) {
        const uint64_t seq = __atomic_load_n(&m->seq, __ATOMIC_ACQUIRE);
        if (seq == want) {
            b->h_st->state = m->state;
            b->h_st->err = m->err;
            b->h_st->nlocal = m->nlocal;
            b->h_st->nghost = m->nghost;
            b->h_st->npairs = m->npairs;
            b->h_st->rebuild_now = m->rebuild_now;
            return 0;
        }
        if (seq > want) return pull_state(b);
        if ((it & 1023u) == 0u) {
            const hipError_t e = hipStreamQuery(b->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(&m->seq, __ATOMIC_ACQUIRE) == want) continue;
                return pull_state(b);
            }
            if (e != hipErrorNotReady) return set_error(SOIL_HIP_EDEVICE, "hipStreamQuery: %s", hipGetErrorString(e));
        }
        __builtin_ia32_pause();
    }
}

int rebuild_tiles(soil_hip_box* b) {
    TileTables tt;
    build_tiles(tt, 0, b->caps);
    const int rc = upload_tiles(tt, b->L, &b->arena_tiles);
    if (rc) return rc;
    b->tiles_bytes = 0;
    for (auto* l : {&tt.box, &tt.local, &tt.local1, &tt.atom, &tt.atom1, &tt.ghost, &tt.cell, &tt.border,
                    &tt.scan_cell, &tt.scan_local, &tt.scan_atom})
        b->tiles_bytes += l->size() * sizeof(Tile);
    b->L.boxes = b->d_ptrs;
    b->L.stream = b->stream;
    b->L.csr_mode = b->h_dev.csr_mode;
    b->L.any_open = !(b->model.periodic[0] && b->model.periodic[1] && b->model.periodic[2]);
    set_transpose_launch(b->L, &b, 1);
    return 0;
}

// Which transpose launches a set of boxes needs (a box without a transpose
// needs none; the launches skip boxes of the other strategy).
void set_transpose_launch(Launch& L, soil_hip_box* const* boxes, size_t n) {
    L.tpull_any = false;
    L.tatomic_any = false;
    for (size_t k = 0; k < n; ++k) {
        const BoxDev& d = boxes[k]->h_dev;
        if (!d.transpose) continue;
        if (d.tpull && d.csr_mode == 5) L.tpull_any = true;
        else L.tatomic_any = true;
    }
}

// The auto transpose strategy after a build: the pull walks about four times
// the CSR entries of each tile's backward rows, which pays while rows are
// short; dense boxes keep the atomics.
int tune_transpose(soil_hip_box* b) {
    if (b->tmode != 2 || !b->tpull_ok || !b->h_dev.transpose || b->h_dev.csr_mode != 5) return 0;
    const uint64_t np = b->h_st->npairs, nl = b->h_st->nlocal;
    int32_t want = b->h_dev.tpull;
    if (np > 4 * nl) want = 0;
    else if (np <= 3 * nl) want = 1;
    if (want == b->h_dev.tpull) return 0;
    b->h_dev.tpull = want;
    // The box's stream is non-blocking, so a copy on the null stream is not
    // ordered after it: this build's transpose kernels (and the step's tail)
    // may still be queued there and must see the strategy they were launched
    // for. Drain the stream first; a switch is rare.
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(&b->d_dev->tpull, &want, sizeof(int32_t), hipMemcpyHostToDevice));
    set_transpose_launch(b->L, &b, 1);
    return 0;
}

// Allocates the atoms arena for caps (copying locals from the previous arena
// when `keep` is set) and the pairs arena.
int alloc_atoms(soil_hip_box* b, const Caps& nc, bool keep) {
    BoxDev nd = b->h_dev;
    Carver dry{nullptr, 0, true};
    carve_atoms(dry, nd, nc, b->cols);
    void* arena = nullptr;
    int rc = alloc_arena(&arena, dry.off + 256);
    if (rc) return rc;
    Carver c{(char*)arena, 0, false};
    carve_atoms(c, nd, nc, b->cols);
    const size_t atoms_bytes = dry.off + 256;
    if (keep) {
        int r2 = pull_desc(b);
        if (r2) return r2;
        r2 = pull_state(b);
        if (r2) return r2;
        const BoxDev& od = b->h_dev;
        const size_t nl = b->nlocal;
        // after a ghost overflow nghost may exceed what the old arena holds
        const size_t n = std::min<size_t>(std::min<size_t>((size_t)b->h_st->nlocal + b->h_st->nghost, nc.atoms),
                                          b->caps.atoms);
        hipError_t ce = hipSuccess;  // the first failed copy, reported after them all
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (bytes && ce == hipSuccess) ce = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b->stream);
        };
        cp(nd.x, od.x, n * 8); cp(nd.y, od.y, n * 8); cp(nd.z, od.z, n * 8);
        cp(nd.vx, od.vx, n * 8); cp(nd.vy, od.vy, n * 8); cp(nd.vz, od.vz, n * 8);
        cp(nd.fx, od.fx, n * 8); cp(nd.fy, od.fy, n * 8); cp(nd.fz, od.fz, n * 8);
        cp(nd.mass, od.mass, n * 8); cp(nd.inv_mass, od.inv_mass, n * 8); cp(nd.rad, od.rad, n * 8);
        cp(nd.tag, od.tag, n * 4); cp(nd.type, od.type, n * 4); cp(nd.origin, od.origin, n * 4);
        cp(nd.imx, od.imx, n * 4); cp(nd.imy, od.imy, n * 4); cp(nd.imz, od.imz, n * 4);
        cp(nd.gswap, od.gswap, n);
        cp(nd.xr, od.xr, n * 32);
        cp(nd.lbx, od.lbx, nl * 8); cp(nd.lby, od.lby, nl * 8); cp(nd.lbz, od.lbz, nl * 8);
        cp(nd.offsets, od.offsets, (nl + 1) * 4);
        cp(nd.tstart, od.tstart, (n + 1) * 4);
        for (size_t i = 0; i < b->cols.size(); ++i)
            cp(nd.cols[i].data, od.cols[i].data, n * b->cols[i].width * 8);
        // the spatial sort's per-cell and per-local arrays (scell, perm, sstart,
        // counts) are not carried over: a build later in this step must bin
        // every atom itself instead of reusing them (same order either way)
        if (b->h_st->local_cells_valid) {
            b->h_st->local_cells_valid = 0;
            HIPCHK(hipMemcpyAsync(&b->d_st->local_cells_valid, &b->h_st->local_cells_valid, 4,
                                  hipMemcpyHostToDevice, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        if (ce != hipSuccess) {
            (void)hipFree(arena);
            return set_error(SOIL_HIP_EDEVICE, "arena copy: %s", hipGetErrorString(ce));
        }
    }
    nd.cap_local = nc.local;
    nd.cap_atoms = nc.atoms;
    nd.cap_cells = nc.cells;
    nd.cap_tiles = nc.tiles;
    if (b->arena_atoms) (void)hipFree(b->arena_atoms);
    b->arena_atoms = arena;
    b->atoms_bytes = atoms_bytes;
    b->h_dev = nd;
    b->caps.local = nc.local;
    b->caps.atoms = nc.atoms;
    b->caps.cells = nc.cells;
    b->caps.tiles = nc.tiles;
    return 0;
}

// Entries per local row that the pair arena keeps through every growth
// (SOIL_HIP_ROW_SLOTS, 4-64, default 12; read at set-up; bit-neutral).
uint32_t row_slots_setting() {
    const char* m = getenv("SOIL_HIP_ROW_SLOTS");
    const int s = m ? atoi(m) : 12;
    return (uint32_t)(s < 4 ? 4 : (s > 64 ? 64 : s));
}

// The pair arena's entries for a request of `cap_pairs` on a box of
// `cap_local` locals: never below row_slots per local (the slots are
// cap_pairs / cap_local), nor 4096.
uint32_t pair_capacity(uint64_t cap_pairs, uint32_t row_slots, uint32_t cap_local) {
    return (uint32_t)std::min<uint64_t>(
        std::max<uint64_t>(std::max<uint64_t>(cap_pairs, 4096), (uint64_t)row_slots * cap_local), 0xFFFFFFF0ull);
}

uint32_t initial_pair_capacity(uint32_t row_slots, uint32_t cap_local) { return pair_capacity(4096, row_slots, cap_local); }

int alloc_pairs(soil_hip_box* b, uint32_t cap_pairs, bool keep) {
    // every growth keeps the row-slot width
    cap_pairs = pair_capacity(cap_pairs, b->row_slots, b->caps.local);
    // the host copy of the descriptor must be current: it is pushed back whole
    // after the growth, and the spatial sort swaps column pointers on the device
    if (b->arena_pairs) {
        const int rp = pull_desc(b);
        if (rp) return rp;
    }
    Caps nc = b->caps;
    nc.pairs = cap_pairs;
    BoxDev nd = b->h_dev;
    Carver dry{nullptr, 0, true};
    carve_pairs(dry, nd, nc);
    void* arena = nullptr;
    int rc = alloc_arena(&arena, dry.off + 256);
    if (rc) return rc;
    Carver c{(char*)arena, 0, false};
    carve_pairs(c, nd, nc);
    if (keep && b->arena_pairs) {
        rc = pull_state(b);
        if (rc) return rc;
        // after an overflow npairs is what the build needed, beyond the old arena
        const size_t np = std::min<size_t>(std::min<size_t>(b->h_st->npairs, cap_pairs), b->caps.pairs);
        if (np) {
            HIPCHK(hipMemcpyAsync(nd.indices, b->h_dev.indices, np * 4, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(nd.tsrc, b->h_dev.tsrc, np * 4, hipMemcpyDeviceToDevice, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    if (b->arena_pairs) (void)hipFree(b->arena_pairs);
    b->arena_pairs = arena;
    b->pairs_bytes = dry.off + 256;
    b->h_dev.indices = nd.indices;
    b->h_dev.ttmp = nd.ttmp;
    b->h_dev.tslot = nd.tslot;
    b->h_dev.tsrc = nd.tsrc;
    b->h_dev.cap_pairs = cap_pairs;
    b->caps.pairs = cap_pairs;
    return 0;
}

int check_box(soil_hip_box* b, bool need_setup) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (need_setup && !b->setup_done) return set_error(SOIL_HIP_ESTATE, "box not set up (call soil_hip_setup)");
    if (b->in_batch) return set_error(SOIL_HIP_ESTATE, "box belongs to a batch; step the batch");
    return 0;
}

int check_err_bits(uint32_t err) {
    if (err & ERR_STENCIL) return set_error(SOIL_HIP_EUNSUPPORTED, "stencil exceeds %d offsets", SOIL_MAX_STENCIL);
    if (err & ERR_REMOVE_PLAN) return set_error(SOIL_HIP_EDEVICE, "atom removal found no kept source for a hole");
    if (err & (ERR_PAIR_CAPACITY | ERR_GHOST_CAPACITY))
        return set_error(SOIL_HIP_ECAPACITY, "device capacity exceeded (err=%u)", err);
    return 0;
}

// Grows the cell arrays when a tilt set after setup needs a larger grid.
int ensure_cells(soil_hip_box* b, uint64_t ncells) {
    if (ncells <= b->caps.cells) return 0;
    Caps nc = b->caps;
    nc.cells = (uint32_t)std::max<uint64_t>(ncells, b->model.max_cells());
    int rc = alloc_atoms(b, nc, true);
    if (rc) return rc;
    rc = push_desc(b);
    if (rc) return rc;
    return rebuild_tiles(b);
}

// Uploads the host model's current geometry into ring slot `ring_pos` (synchronous).
int upload_geom_now(soil_hip_box* b) {
    Geom& g = b->h_ring[b->ring_pos];
    if (b->model.fill(g)) return set_error(SOIL_HIP_EUNSUPPORTED, "stencil exceeds %d offsets", SOIL_MAX_STENCIL);
    int rcc = ensure_cells(b, g.ncells);
    if (rcc) return rcc;
    b->last_geom = g;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(b->d_ring + b->ring_pos, &g, sizeof(Geom), hipMemcpyHostToDevice));
    b->ring_ok[b->ring_pos] = 1;
    return 0;
}

void resolve_timing(soil_hip_box* b) {
    std::vector<soil_hip_box::Pending> keep;
    for (auto& p : b->pending) {
        if (hipEventQuery(p.e2) == hipSuccess) {
            float ms_all = 0.f, ms_fill = 0.f, ms_count = 0.f;
            (void)hipEventElapsedTime(&ms_all, p.e0, p.e2);
            (void)hipEventElapsedTime(&ms_fill, p.e1, p.e2);
            (void)hipEventElapsedTime(&ms_count, p.e0, p.ec);
            b->t_builds += 1.0;
            b->t_csr_ms += ms_all;
            b->t_fill_ms += ms_fill;
            b->t_count_ms += ms_count;
            b->ev_pool.push_back(p.e0);
            b->ev_pool.push_back(p.ec);
            b->ev_pool.push_back(p.e1);
            b->ev_pool.push_back(p.e2);
        } else {
            keep.push_back(p);
        }
    }
    b->pending.swap(keep);
}

hipEvent_t take_event(soil_hip_box* b) {
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    // timing only: no system-scope fence at record (it would write back caches
    // in the middle of the step being timed)
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;  // that build goes untimed
    return e;
}

// borders, forward comm and the neighbor build of a rebuild step, with one
// state read at the end: a ghost overflow makes every build kernel skip, so
// the host grows the atoms arena and redoes the three; a pair overflow skips
// the fill, and the host grows the pairs and redoes the build.
// The rebuild's borders, build and publish, enqueued without waiting; the
// timing events (when timing) go to `ev`.
static void rebuild_launch(soil_hip_box* b, hipEvent_t ev[4]) {
    const bool timed = b->timing && (b->timing_ctr++ % b->timing) == 0;
    for (int k = 0; k < 4; ++k) ev[k] = timed ? take_event(b) : nullptr;
    if (timed && !(ev[0] && ev[1] && ev[2] && ev[3])) {  // an event could not be made: this build goes untimed
        for (int k = 0; k < 4; ++k)
            if (ev[k]) b->ev_pool.push_back(ev[k]);
        for (int k = 0; k < 4; ++k) ev[k] = nullptr;
    }
    // borders, ghosts, build (count, scan flags the pair capacity, fill), publish
    phase_rebuild_fused(b->L, ev[0], ev[1], ev[2], ev[3], true);
    b->pub_expect += 1;
}

// Waits for rebuild_launch's publish. A build that ran out of ghost or pair
// capacity grows the arena and is redone (`redone`); the kernels enqueued
// after the publish skip their work while the error bits are set.
static int rebuild_check(soil_hip_box* b, hipEvent_t ev[4], bool& redone) {
    redone = false;
    int rc = wait_mirror(b);
    if (rc) return rc;
    const bool grow = (b->h_st->err & (ERR_GHOST_CAPACITY | ERR_PAIR_CAPACITY)) != 0;
    if (grow) {
        rc = pull_state(b);  // drains the skipped kernels; every field current
        if (rc) return rc;
    }
    for (int attempt = 0; (b->h_st->err & ERR_GHOST_CAPACITY) && attempt < 4; ++attempt) {
        Caps nc = b->caps;
        uint64_t ghost_cap = (uint64_t)(b->caps.atoms - b->caps.local);
        ghost_cap = ghost_cap * 2 + 4096;
        nc.atoms = (uint32_t)(b->caps.local + ghost_cap);
        nc.tiles = ceil_div(nc.atoms, 64);
        rc = alloc_atoms(b, nc, true);
        if (rc) return rc;
        b->h_st->err &= ~ERR_GHOST_CAPACITY;
        HIPCHK(hipMemcpy(&b->d_st->err, &b->h_st->err, 4, hipMemcpyHostToDevice));
        rc = push_desc(b);
        if (!rc) rc = rebuild_tiles(b);
        if (rc) return rc;
        phase_borders(b->L);
        phase_forward(b->L);
        phase_build_fused(b->L, nullptr, nullptr, nullptr, nullptr);
        rc = pull_state(b);
        if (rc) return rc;
        redone = true;
    }
    if (b->h_st->err & ERR_GHOST_CAPACITY) return set_error(SOIL_HIP_ECAPACITY, "ghost capacity growth failed");
    if (b->h_st->err & ERR_PAIR_CAPACITY) {
        // grow pairs and redo the build (the state is intact up to the fill)
        uint32_t np = b->h_st->npairs;
        rc = alloc_pairs(b, (uint32_t)std::min<uint64_t>((uint64_t)np + np / 2 + 4096, 0xFFFFFFF0u), false);
        if (rc) return rc;
        b->h_st->err &= ~ERR_PAIR_CAPACITY;
        rc = push_desc(b);
        if (!rc) rc = rebuild_tiles(b);
        if (rc) return rc;
        // rebuilds was counted once by build_prep; the redo re-runs the whole build,
        // so compensate the counter first
        b->h_st->rebuilds -= 1;
        b->h_st->state = 0;  // build_prep flips to CommunicateOnly again
        HIPCHK(hipMemcpy(b->d_st, b->h_st, sizeof(BoxState), hipMemcpyHostToDevice));
        phase_build(b->L, nullptr, nullptr, nullptr, nullptr);
        rc = pull_state(b);
        if (rc) return rc;
        if (b->h_st->err & ERR_PAIR_CAPACITY) return set_error(SOIL_HIP_ECAPACITY, "pair capacity growth failed");
        redone = true;
    }
    if (ev[0]) {
        if (redone || !b->h_st->rebuild_now) {  // the timed build was skipped or redone: no sample
            for (int k = 0; k < 4; ++k) b->ev_pool.push_back(ev[k]);
        } else {
            b->pending.push_back({ev[0], ev[1], ev[2], ev[3]});
        }
    }
    return 0;
}

int step_plain_loop(soil_hip_box* b, int nsteps, bool& pending_final);

int step_plain(soil_hip_box* b, int nsteps) {
    int rc0 = pull_state(b);  // exact publish count (batches and other calls publish too)
    if (rc0) return rc0;
    // final_integration of the previous step, when it is carried into this
    // step's initial_integration kernel; always launched before returning
    bool pending_final = false;
    int rc = step_plain_loop(b, nsteps, pending_final);
    if (pending_final) phase_final(b->L);
    if (rc) return rc;
    if (b->timing) resolve_timing(b);
    return 0;
}

// Stages the host model's geometry in this step's ring slot (the stream is
// ordered, so a slot rewritten later in the step reaches the later kernels).
// A slot that already holds the same bytes is not uploaded again (a box whose
// geometry does not change reuses it every ring_len steps).
static int stage_geom(soil_hip_box* b, Geom& g) {
    Geom t;
    if (b->model.fill(t)) return set_error(SOIL_HIP_EUNSUPPORTED, "stencil exceeds %d offsets", SOIL_MAX_STENCIL);
    if (t.ncells > b->caps.cells) {
        int rcc = ensure_cells(b, t.ncells);
        if (rcc) return rcc;
    }
    b->last_geom = t;
    uint8_t& ok = b->ring_ok[b->ring_pos];
    if (ok && std::memcmp(&g, &t, sizeof(Geom)) == 0) return 0;
    std::memcpy(&g, &t, sizeof(Geom));  // fill() cleared the padding, so equal geometries compare equal
    HIPCHK(hipMemcpyAsync(b->d_ring + b->ring_pos, &g, sizeof(Geom), hipMemcpyHostToDevice, b->stream));
    ok = 1;
    return 0;
}

static double ext_value(uint64_t key) {
    const uint64_t u = (key >> 63) ? (key & 0x7FFFFFFFFFFFFFFFull) : ~key;
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// shrink_wrap (domain.rs:509-569) for this step: the locals' extents after the
// kick-drift come back to the host, which moves the bounds; sort_atoms_by_bin's
// bookkeeping then sees the change (neighbor.rs:836-841).
static int shrink_wrap_now(soil_hip_box* b) {
    phase_extent(b->L);
    int rc = pull_state(b);
    if (rc) return rc;
    double pmin[3], pmax[3];
    for (int k = 0; k < 3; ++k) {
        pmin[k] = ext_value(b->h_st->ext_key[k]);
        pmax[k] = ext_value(b->h_st->ext_key[3 + k]);
    }
    b->model.shrink_wrap_update(pmin, pmax, b->h_st->nlocal);
    return 0;
}

int step_plain_loop(soil_hip_box* b, int nsteps, bool& pending_final) {
    const bool shrink = b->model.has_shrink();
    for (int s = 0; s < nsteps; ++s) {
        if (!b->model.apply_deform())
            return set_error(SOIL_HIP_EINVAL, "[deform] an axis collapsed to zero size (soil_deform/src/lib.rs:506-515)");
        // a shrink-wrap box learns its bounds after the kick-drift: its sort
        // bookkeeping waits until then, and the kick reads the bounds so far
        if (!shrink) b->model.sort_prepare();
        Geom& g = b->h_ring[b->ring_pos];
        int rc = stage_geom(b, g);
        if (rc) return rc;
        // the step's deform work (remap, one-time shear profile) runs inside the kick-drift
        if (shrink) {
            phase_kick(b->L, pending_final);
            rc = shrink_wrap_now(b);
            if (rc) return rc;
            b->model.sort_prepare();
            rc = stage_geom(b, g);  // may regrid and regrow the cell arrays
            if (rc) return rc;
            phase_zero_counts(b->L);  // k_initial zeroed the counts of the provisional grid only
            phase_decide_fused(b->L);
        } else if (pending_final) {
            phase_final_initial_decide(b->L);
        } else {
            phase_initial_decide(b->L);
        }
        b->pub_expect += 1;  // k_decide publishes the decision
        pending_final = false;
        // A box whose step before last rebuilt enqueues the rebuild path
        // without waiting for this step's decision (rebuilds come in runs, or
        // every other step): each kernel on it acts on the device's decision
        // (pbc on FullRebuild; the sorts, borders and build on rebuild_now;
        // the forward refresh otherwise), so a step that does not rebuild runs
        // it as the forward step.
        const bool ahead = !shrink && !b->hook && (b->rebuild_hist & 2u) != 0;
        bool rebuild = true;
        if (!ahead) {
            rc = wait_mirror(b);
            if (rc) return rc;
            if (b->h_st->err) return check_err_bits(b->h_st->err);
            rebuild = b->h_st->state == 0 || g.periodic_sort || g.bounds_changed;
        }
        bool zero_cols = false;
        for (const auto& c : b->cols) zero_cols |= (c.flags & SOIL_HIP_COL_ZERO) != 0;
        if (!b->consumer || zero_cols) phase_zero_fused(b->L);
        // consumer, hook, reverse comm and the step's end
        auto tail = [&]() -> int {
            if (b->consumer) phase_force_fused(b->L);
            if (b->hook) {
                // the host counts are current (read after the decision, the
                // borders and the build); the column buffers swap at each sort
                int r = pull_desc(b);
                if (!r) r = pull_state(b);
                if (!r) r = call_hook(b, true, b->h_st->steps);
                if (r) return r;
            }
            if (b->model.newton) {
                for (int d = 0; d < 3; ++d) b->L.rev_split[d] = g.disjoint[d] ? 0 : 1;
                phase_reverse(b->L);
            }
            phase_step_end(b->L);
            return 0;
        };
        if (rebuild) {
            phase_pbc_sort_fused(b->L);
            hipEvent_t ev[4];
            rebuild_launch(b, ev);
            // the tail goes out before the host reads the build's counts
            // (a hook reads them first); after a capacity error it was a no-op
            const bool early = !b->hook;
            if (early) {
                rc = tail();
                if (rc) return rc;
            }
            bool redone = false;
            rc = rebuild_check(b, ev, redone);
            if (rc) return rc;
            rc = tune_transpose(b);
            if (rc) return rc;
            if (!early || redone) {
                rc = tail();
                if (rc) return rc;
            }
            b->rebuild_hist = ((b->rebuild_hist << 1) | (b->h_st->rebuild_now ? 1u : 0u)) & 3u;
            pending_final = true;  // launched with the next step's initial integration, or on return
            if (b->h_st->err) return check_err_bits(b->h_st->err);
        } else {
            phase_forward(b->L);  // k_decide<true> left rebuild_now = 0 and the sort flags cleared
            rc = tail();
            if (rc) return rc;
            b->rebuild_hist = (b->rebuild_hist << 1) & 3u;
            pending_final = true;
        }
        b->ring_pos = (b->ring_pos + 1) % b->ring_len;
    }
    return 0;
}

}  // namespace soil_hip

namespace soil_hip {
// Device bytes a box will hold after setup (its arenas at their initial capacities).
size_t box_bytes_estimate(soil_hip_box* b) {
    HostModel m = b->model;
    m.setup(b->hrad.data(), b->nlocal);
    Caps c;
    c.local = (uint32_t)b->nlocal;
    const uint64_t ghost = m.ghost_estimate(b->nlocal);
    c.atoms = (uint32_t)std::min<uint64_t>(b->nlocal + ghost + ghost / 2 + 1024, 0x7FFFFFFFull);
    c.cells = (uint32_t)m.max_cells();
    c.tiles = ceil_div(c.atoms, 64);
    c.pairs = initial_pair_capacity(row_slots_setting(), c.local);
    BoxDev d{};
    Carver a{nullptr, 0, true};
    carve_atoms(a, d, c, b->cols);
    Carver p{nullptr, 0, true};
    carve_pairs(p, d, c);
    TileTables tt;
    build_tiles(tt, 0, c);
    const std::vector<Tile>* lists[11] = {&tt.box,  &tt.local,  &tt.local1,    &tt.atom,       &tt.atom1,   &tt.ghost,
                                          &tt.cell, &tt.border, &tt.scan_cell, &tt.scan_local, &tt.scan_atom};
    size_t tiles = 0;
    for (auto* l : lists) tiles += l->size();
    // arenas as alloc_atoms / alloc_pairs size them, the tile tables, the
    // geometry ring, and the descriptor, state and view records
    return (a.off + 256) + (p.off + 256) + tiles * sizeof(Tile) + sizeof(Geom) * b->ring_len + 4096;
}

size_t box_bytes_allocated(soil_hip_box* b) {
    if (!b->setup_done) return 0;
    return b->atoms_bytes + b->pairs_bytes + b->tiles_bytes + sizeof(Geom) * b->ring_len + sizeof(BoxDev) +
           sizeof(BoxState) + sizeof(soil_hip_device_view) + sizeof(BoxDev*);
}
}  // namespace soil_hip

// ─────────────────────────────── C ABI ─────────────────────────────────────
extern "C" {

const char* soil_hip_last_error(void) { return g_last_error.c_str(); }

int soil_hip_device_count(int* out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (out) *out = n;
    return 0;
}

int soil_hip_ctx_create(int device, soil_hip_ctx** out) {
    if (!out) return set_error(SOIL_HIP_EINVAL, "null out");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(SOIL_HIP_EDEVICE, "no HIP device");
    if (device < 0 || device >= n) return set_error(SOIL_HIP_EINVAL, "device %d out of range (%d)", device, n);
    HIPCHK(hipSetDevice(device));
    // the allocation-free entry points' workspace on this device
    if (scratch_init(device) != 0)
        return set_error(SOIL_HIP_EDEVICE, "hipGetSymbolAddress of the fixed workspace failed on device %d", device);
    soil_hip_ctx* c = new soil_hip_ctx();
    c->device = device;
    *out = c;
    return 0;
}

int soil_hip_ctx_destroy(soil_hip_ctx* ctx) {
    delete ctx;
    return 0;
}

int soil_hip_box_create(soil_hip_ctx* ctx, const soil_hip_box_desc* desc, soil_hip_box** out) {
    if (!ctx || !desc || !out) return set_error(SOIL_HIP_EINVAL, "null argument");
    for (int k = 0; k < 3; ++k)
        if (!(desc->hi[k] > desc->lo[k])) return set_error(SOIL_HIP_EINVAL, "box hi must exceed lo on every axis");
    if (!(desc->skin_fraction >= 1.0)) return set_error(SOIL_HIP_EINVAL, "skin_fraction must be >= 1");
    if (!(desc->dt > 0.0)) return set_error(SOIL_HIP_EINVAL, "dt must be > 0");
    if (desc->has_shear && (!desc->periodic[0] || !desc->periodic[1]))
        return set_error(SOIL_HIP_EINVAL, "[deform] xy shear requires periodic x and y boundaries");
    HIPCHK(hipSetDevice(ctx->device));
    soil_hip_box* b = new soil_hip_box();
    b->ctx = ctx;
    b->desc = *desc;
    b->model.init(*desc);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return set_error(SOIL_HIP_EDEVICE, "hipStreamCreate failed");
    }
    *out = b;
    return 0;
}

int soil_hip_box_destroy(soil_hip_box* b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& p : b->pending) {
        b->ev_pool.push_back(p.e0);
        b->ev_pool.push_back(p.e1);
        b->ev_pool.push_back(p.e2);
    }
    for (auto e : b->ev_pool) (void)hipEventDestroy(e);
    if (b->arena_atoms) (void)hipFree(b->arena_atoms);
    if (b->arena_pairs) (void)hipFree(b->arena_pairs);
    if (b->arena_tiles) (void)hipFree(b->arena_tiles);
    if (b->d_dev) (void)hipFree(b->d_dev);
    if (b->d_ptrs) (void)hipFree(b->d_ptrs);
    if (b->d_st) (void)hipFree(b->d_st);
    if (b->d_view) (void)hipFree(b->d_view);
    if (b->d_ring) (void)hipFree(b->d_ring);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_mirror) (void)hipHostFree(b->h_mirror);
    if (b->h_ring) (void)hipHostFree(b->h_ring);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int soil_hip_register_column(soil_hip_box* b, int width, uint32_t flags, int* col_id) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "register columns before setup");
    if (width < 1 || width > 64) return set_error(SOIL_HIP_EINVAL, "column width must be 1..64");
    if (flags & ~7u) return set_error(SOIL_HIP_EINVAL, "unknown column flags");
    if (b->cols.size() >= SOIL_MAX_COLS) return set_error(SOIL_HIP_EINVAL, "at most %d columns", SOIL_MAX_COLS);
    HostCol c;
    c.width = width;
    c.flags = flags;
    c.data.assign(b->nlocal * (size_t)width, 0.0);
    b->cols.push_back(c);
    if (col_id) *col_id = (int)b->cols.size() - 1;
    return 0;
}

int soil_hip_upload(soil_hip_box* b, size_t n, const double* pos, const double* vel, const double* force,
                    const double* mass, const double* radius, const uint32_t* tag, const uint32_t* type,
                    const int32_t* image) {
    if (!b || !pos || !mass || !radius) return set_error(SOIL_HIP_EINVAL, "null argument");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "upload before setup");
    if (b->nlocal + n > 0x7FFFFFFFull) return set_error(SOIL_HIP_EINVAL, "too many atoms");
    for (size_t i = 0; i < n; ++i) {
        if (!(mass[i] > 0.0) || !(radius[i] > 0.0)) return set_error(SOIL_HIP_EINVAL, "mass and radius must be > 0");
        b->hx.push_back(pos[3 * i]);
        b->hy.push_back(pos[3 * i + 1]);
        b->hz.push_back(pos[3 * i + 2]);
        b->hvx.push_back(vel ? vel[3 * i] : 0.0);
        b->hvy.push_back(vel ? vel[3 * i + 1] : 0.0);
        b->hvz.push_back(vel ? vel[3 * i + 2] : 0.0);
        b->hfx.push_back(force ? force[3 * i] : 0.0);
        b->hfy.push_back(force ? force[3 * i + 1] : 0.0);
        b->hfz.push_back(force ? force[3 * i + 2] : 0.0);
        b->hmass.push_back(mass[i]);
        b->hrad.push_back(radius[i]);
        b->htag.push_back(tag ? tag[i] : (uint32_t)(b->nlocal + i));
        b->htype.push_back(type ? type[i] : 0u);
        b->himx.push_back(image ? image[3 * i] : 0);
        b->himy.push_back(image ? image[3 * i + 1] : 0);
        b->himz.push_back(image ? image[3 * i + 2] : 0);
    }
    b->nlocal += n;
    for (auto& c : b->cols) c.data.resize(b->nlocal * (size_t)c.width, 0.0);
    return 0;
}

int soil_hip_upload_column(soil_hip_box* b, int col, const double* data, size_t n) {
    if (!b || !data || col < 0 || col >= (int)b->cols.size()) return set_error(SOIL_HIP_EINVAL, "bad column");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "upload columns before setup");
    if (n > b->nlocal) return set_error(SOIL_HIP_EINVAL, "more rows than atoms");
    std::memcpy(b->cols[col].data.data(), data, n * b->cols[col].width * sizeof(double));
    return 0;
}

int soil_hip_setup(soil_hip_box* b) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "setup already done");
    HIPCHK(hipSetDevice(b->ctx->device));
    const size_t n = b->nlocal;
    b->est_bytes = box_bytes_estimate(b);
    b->model.setup(b->hrad.data(), n);
    if (b->model.max_stencil() > SOIL_MAX_STENCIL)
        return set_error(SOIL_HIP_EUNSUPPORTED, "stencil of %zu offsets exceeds %d", b->model.max_stencil(),
                         SOIL_MAX_STENCIL);
    Caps c;
    c.local = (uint32_t)n;
    const uint64_t ghost = b->model.ghost_estimate(n);
    c.atoms = (uint32_t)std::min<uint64_t>(n + ghost + ghost / 2 + 1024, 0x7FFFFFFFull);
    c.cells = (uint32_t)b->model.max_cells();
    c.tiles = ceil_div(c.atoms, 64);
    b->h_dev = BoxDev{};
    b->h_dev.ncols = (int32_t)b->cols.size();
    for (size_t i = 0; i < b->cols.size(); ++i) {
        b->h_dev.cols[i].width = b->cols[i].width;
        b->h_dev.cols[i].flags = b->cols[i].flags;
    }
    HIPCHK(hipMalloc(&b->d_dev, sizeof(BoxDev)));
    HIPCHK(hipMalloc(&b->d_ptrs, sizeof(BoxDev*)));
    HIPCHK(hipMemcpy(b->d_ptrs, &b->d_dev, sizeof(BoxDev*), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&b->d_st, sizeof(BoxState)));
    HIPCHK(hipHostMalloc(&b->h_st, sizeof(BoxState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&b->h_mirror, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(b->h_mirror, 0, sizeof(HostMirror));
    void* mirror_dev = nullptr;
    HIPCHK(hipHostGetDevicePointer(&mirror_dev, b->h_mirror, 0));
    HIPCHK(hipMalloc(&b->d_view, sizeof(soil_hip_device_view)));
    HIPCHK(hipMalloc(&b->d_ring, sizeof(Geom) * b->ring_len));
    HIPCHK(hipHostMalloc(&b->h_ring, sizeof(Geom) * b->ring_len, hipHostMallocDefault));
    b->ring_ok.assign(b->ring_len, 0);
    int rc = alloc_atoms(b, c, false);
    if (rc) return rc;
    b->row_slots = row_slots_setting();
    rc = alloc_pairs(b, 4096, false);  // raised to row_slots per local
    if (rc) return rc;
    BoxDev& d = b->h_dev;
    // locals (the primary buffers are the ones named without the a- prefix)
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    std::vector<double> inv(n);
    for (size_t i = 0; i < n; ++i) inv[i] = 1.0 / b->hmass[i];
    std::vector<int32_t> zeros_i(n, 0);
    HIPCHK(up(d.x, b->hx.data(), n * 8)); HIPCHK(up(d.y, b->hy.data(), n * 8)); HIPCHK(up(d.z, b->hz.data(), n * 8));
    HIPCHK(up(d.vx, b->hvx.data(), n * 8)); HIPCHK(up(d.vy, b->hvy.data(), n * 8)); HIPCHK(up(d.vz, b->hvz.data(), n * 8));
    HIPCHK(up(d.fx, b->hfx.data(), n * 8)); HIPCHK(up(d.fy, b->hfy.data(), n * 8)); HIPCHK(up(d.fz, b->hfz.data(), n * 8));
    HIPCHK(up(d.mass, b->hmass.data(), n * 8)); HIPCHK(up(d.inv_mass, inv.data(), n * 8));
    HIPCHK(up(d.rad, b->hrad.data(), n * 8));
    HIPCHK(up(d.tag, b->htag.data(), n * 4)); HIPCHK(up(d.type, b->htype.data(), n * 4));
    HIPCHK(up(d.origin, zeros_i.data(), n * 4));
    HIPCHK(up(d.imx, b->himx.data(), n * 4)); HIPCHK(up(d.imy, b->himy.data(), n * 4)); HIPCHK(up(d.imz, b->himz.data(), n * 4));
    for (size_t i = 0; i < b->cols.size(); ++i)
        HIPCHK(up(d.cols[i].data, b->cols[i].data.data(), n * b->cols[i].width * 8));
    {
        std::vector<double> xr(4 * n);
        for (size_t i = 0; i < n; ++i) {
            xr[4 * i] = b->hx[i];
            xr[4 * i + 1] = b->hy[i];
            xr[4 * i + 2] = b->hz[i];
            xr[4 * i + 3] = b->hrad[i];
        }
        HIPCHK(up(d.xr, xr.data(), n * 32));
    }
    d.st = b->d_st;
    d.view = b->d_view;
    d.mirror = (HostMirror*)mirror_dev;
    d.geom_ring = b->d_ring;
    d.ring_len = b->ring_len;
    d.ring_stride = 1;
    d.harmonic_k = b->harmonic_k;
    b->max_radius = 0.0;
    for (size_t i = 0; i < n; ++i) b->max_radius = std::fmax(b->max_radius, b->hrad[i]);
    d.fix_fast = fix_fast_ok(b->harmonic_k, b->max_radius);
    d.dt = b->model.dt;
    d.consumer = b->consumer ? 1 : 0;
    d.transpose = (b->consumer || (b->hook && (b->hook_flags & SOIL_HIP_HOOK_TRANSPOSE))) ? 1 : 0;
    {
        const char* m = getenv("SOIL_HIP_CSR_MODE");
        d.csr_mode = m ? atoi(m) : 5;
        if (d.csr_mode < 0 || d.csr_mode > 5) d.csr_mode = 5;
    }
    {
        // transpose strategy (bit-neutral): atomic (default), pull or auto;
        // auto starts from the expected pairs per row of a uniform box,
        // 4 phi sf^3, and follows each build's count (a clustered box switches
        // after its first). The pull is opt-in until it has run on the device
        // with its host-side fixes.
        const char* m = getenv("SOIL_HIP_TRANSPOSE");
        b->tmode = 0;
        if (m && (!strcmp(m, "pull") || !strcmp(m, "1"))) b->tmode = 1;
        if (m && !strcmp(m, "auto")) b->tmode = 2;
        b->tpull_ok = b->model.max_stencil() + 1 <= (size_t)TP_MAXR;
        double vol = 1.0, vs = 0.0;
        for (int k = 0; k < 3; ++k) vol *= b->model.hi[k] - b->model.lo[k];
        for (size_t i = 0; i < n; ++i) vs += b->hrad[i] * b->hrad[i] * b->hrad[i];
        const double phi = vol > 0.0 ? 4.18879020478639 * vs / vol : 1.0;
        const double sf = b->model.skin_fraction;
        const bool sparse = 4.0 * phi * sf * sf * sf <= 3.0;
        d.tpull = (b->tpull_ok && (b->tmode == 1 || (b->tmode == 2 && sparse))) ? 1 : 0;
    }
    BoxState st{};
    st.nlocal = (uint32_t)n;
    st.state = 0;  // CurrentState(FullRebuild), neighbor.rs:545
    st.ring_pos = 0;
    st.step_limit = ~0ull;
    std::memcpy(b->h_st, &st, sizeof(st));
    HIPCHK(hipMemcpy(b->d_st, &st, sizeof(BoxState), hipMemcpyHostToDevice));
    b->ring_pos = 0;
    {
        // blocks per launch at most (tiles beyond it loop); bit-neutral tuning knob
        const char* m = getenv("SOIL_HIP_MAX_GRID");
        const int g = m ? atoi(m) : 4096;
        b->L.maxgrid = (uint32_t)(g < 256 ? 256 : (g > 65536 ? 65536 : g));
    }
    rc = push_desc(b);
    if (!rc) rc = rebuild_tiles(b);
    if (rc) return rc;
    b->setup_done = true;
    rc = upload_geom_now(b);
    if (rc) return rc;
    // every slot starts with the set-up geometry: a box whose geometry stays
    // the same then uploads none while it steps
    for (uint32_t k = 1; k < b->ring_len; ++k) std::memcpy(&b->h_ring[k], &b->h_ring[0], sizeof(Geom));
    HIPCHK(hipMemcpy(b->d_ring, b->h_ring, sizeof(Geom) * b->ring_len, hipMemcpyHostToDevice));
    b->ring_ok.assign(b->ring_len, 1);
    // host copies no longer needed
    b->hx.clear(); b->hy.clear(); b->hz.clear(); b->hvx.clear(); b->hvy.clear(); b->hvz.clear();
    b->hfx.clear(); b->hfy.clear(); b->hfz.clear(); b->hmass.clear(); b->htag.clear(); b->htype.clear();
    b->himx.clear(); b->himy.clear(); b->himz.clear();
    return 0;
}

int soil_hip_set_boundary(soil_hip_box* b, int dim, int type) {
    if (!b || dim < 0 || dim > 2 || type < 0 || type > 2) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "set boundaries before setup");
    if (type != 0 && dim < 2 && b->model.has_shear)
        return set_error(SOIL_HIP_EINVAL, "[deform] xy shear requires periodic x and y boundaries");
    b->model.periodic[dim] = type == 0;
    b->model.shrink[dim] = type == 2;
    return 0;
}

int soil_hip_set_shrink_wrap_padding(soil_hip_box* b, double padding) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    b->model.shrink_padding = padding;
    return 0;
}

int soil_hip_set_axis_deform(soil_hip_box* b, int dim, int style, double value, double target_lo, double target_hi) {
    if (!b || dim < 0 || dim > 2 || style < -1 || style > 2) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    if (b->setup_done) return set_error(SOIL_HIP_ESTATE, "define deformation before setup (setup_deform captures bounds)");
    if (style == 2 && !(target_hi > target_lo)) return set_error(SOIL_HIP_EINVAL, "final style needs lo < hi");
    b->model.axis_style[dim] = style;
    b->model.axis_value[dim] = value;
    b->model.axis_target_lo[dim] = target_lo;
    b->model.axis_target_hi[dim] = target_hi;
    return 0;
}

int soil_hip_set_deform_remap(soil_hip_box* b, int remap) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    b->model.remap = remap != 0;
    return 0;
}

int soil_hip_set_stage_steps(soil_hip_box* b, uint64_t nsteps) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    b->model.stage_steps = nsteps;
    return 0;
}

int soil_hip_get_domain(soil_hip_box* b, double* out9) {
    if (!b || !out9) return set_error(SOIL_HIP_EINVAL, "null argument");
    for (int k = 0; k < 3; ++k) {
        out9[k] = b->model.lo[k];
        out9[3 + k] = b->model.hi[k];
        out9[6 + k] = b->model.size[k];
    }
    return 0;
}

int soil_hip_set_consumer(soil_hip_box* b, int enabled, double k) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (b->setup_done) {
        int rc = pull_state(b);
        if (rc) return rc;
        if (b->h_st->rebuilds != 0)
            return set_error(SOIL_HIP_ESTATE, "enable the consumer before the first neighbor build");
    }
    b->consumer = enabled != 0;
    b->harmonic_k = k;
    if (b->setup_done) {
        b->h_dev.consumer = b->consumer ? 1 : 0;
        b->h_dev.harmonic_k = k;
        b->h_dev.fix_fast = fix_fast_ok(k, b->max_radius);
        b->h_dev.transpose = (b->consumer || (b->hook && (b->hook_flags & SOIL_HIP_HOOK_TRANSPOSE))) ? 1 : 0;
        // patch the fields in place (the rest of the device descriptor may hold swapped pointers)
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipMemcpy(&b->d_dev->harmonic_k, &b->h_dev.harmonic_k, sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&b->d_dev->consumer, &b->h_dev.consumer, sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&b->d_dev->fix_fast, &b->h_dev.fix_fast, sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&b->d_dev->transpose, &b->h_dev.transpose, sizeof(int32_t), hipMemcpyHostToDevice));
    }
    return 0;
}

int soil_hip_set_force_hook(soil_hip_box* b, soil_hip_force_hook fn, void* user, int flags) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (b->in_batch) return set_error(SOIL_HIP_ESTATE, "set hooks before batching the box");
    const bool want_t = fn && (flags & SOIL_HIP_HOOK_TRANSPOSE);
    if (b->setup_done && want_t && !b->h_dev.transpose) {
        int rc = pull_state(b);
        if (rc) return rc;
        if (b->h_st->rebuilds != 0)
            return set_error(SOIL_HIP_ESTATE, "ask for the transpose before the first neighbor build");
    }
    b->hook = fn;
    b->hook_user = user;
    b->hook_flags = fn ? flags : 0;
    if (b->setup_done) {
        b->h_dev.transpose = (b->consumer || want_t) ? 1 : 0;
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipMemcpy(&b->d_dev->transpose, &b->h_dev.transpose, sizeof(int32_t), hipMemcpyHostToDevice));
    }
    return 0;
}

}  // extern "C"

namespace soil_hip {
void fill_view(soil_hip_box* b, soil_hip_device_view& v, bool exact, uint64_t step) {
    std::memset(&v, 0, sizeof(v));
    const BoxDev& d = b->h_dev;
    v.cap_local = d.cap_local;
    v.cap_atoms = d.cap_atoms;
    v.cap_pairs = d.cap_pairs;
    v.step = step;
    if (!exact) {  // inside a batch: counts are the capacities, pointers come from d_view
        v.nlocal = d.cap_local;
        v.nghost = d.cap_atoms - d.cap_local;
        v.npairs = d.cap_pairs;
        return;
    }
    v.nlocal = b->h_st->nlocal;
    v.nghost = b->h_st->nghost;
    v.npairs = b->h_st->npairs;
    v.rebuilt = b->h_st->rebuild_now;
    v.ncols = (uint32_t)d.ncols;
    v.x = d.x; v.y = d.y; v.z = d.z;
    v.vx = d.vx; v.vy = d.vy; v.vz = d.vz;
    v.fx = d.fx; v.fy = d.fy; v.fz = d.fz;
    v.xr = d.xr;
    v.mass = d.mass; v.inv_mass = d.inv_mass; v.radius = d.rad;
    v.tag = d.tag; v.type = d.type; v.origin = d.origin;
    v.offsets = d.offsets;
    v.indices = d.indices;
    v.tstart = d.transpose ? d.tstart : nullptr;
    v.tsrc = d.transpose ? d.tsrc : nullptr;
    for (int c = 0; c < d.ncols && c < SOIL_HIP_MAX_COLS; ++c) {
        v.cols[c] = d.cols[c].data;
        v.col_width[c] = d.cols[c].width;
    }
}

// Refreshes the device view on the stream, then calls the hook (which
// enqueues its kernels behind the refresh).
int call_hook(soil_hip_box* b, bool exact, uint64_t step) {
    phase_sync_view(b->L);
    soil_hip_device_view v;
    fill_view(b, v, exact, step);
    const int rc = b->hook(b->hook_user, &v, b->h_dev.view, (void*)b->stream);
    if (rc) return set_error(SOIL_HIP_EHOOK, "force hook returned %d", rc);
    return 0;
}
}  // namespace soil_hip

extern "C" {

int soil_hip_get_device_view(soil_hip_box* b, soil_hip_device_view* out) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (!out) return set_error(SOIL_HIP_EINVAL, "null view");
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);  // column buffers swap with every spatial sort
    if (rc) return rc;
    fill_view(b, *out, true, b->h_st->steps);
    return 0;
}


// ── per-phase systems ──────────────────────────────────────────────────────
#define PHASE_PROLOGUE(b)              \
    do {                               \
        int rc_ = check_box(b, true);  \
        if (rc_) return rc_;           \
    } while (0)

int soil_hip_apply_deform(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    if (!b->model.apply_deform())
        return set_error(SOIL_HIP_EINVAL, "[deform] an axis collapsed to zero size (soil_deform/src/lib.rs:506-515)");
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_deform(b->L);
    return 0;
}

int soil_hip_initial_integration(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_initial(b->L);
    return 0;
}

int soil_hip_decide_rebuild(soil_hip_box* b, int* full_rebuild) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    // decide_rebuild removes nothing: keep ghosts until remove_ghost_atoms
    phase_decide_only(b->L);
    rc = pull_state(b);
    if (rc) return rc;
    if (full_rebuild) *full_rebuild = b->h_st->state == 0 ? 1 : 0;
    return 0;
}

int soil_hip_shrink_wrap(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    if (!b->model.has_shrink()) return 0;
    return shrink_wrap_now(b);
}

int soil_hip_remove_ghost_atoms(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    phase_remove_ghosts(b->L);
    return 0;
}

int soil_hip_zero_all_forces(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    phase_zero(b->L);
    return 0;
}

int soil_hip_pbc(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_pbc(b->L);
    return 0;
}

int soil_hip_sort_atoms_by_bin(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    b->model.sort_prepare();
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_sort(b->L);
    return 0;
}

int soil_hip_borders(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_set_rebuild_now(b->L);
    phase_borders(b->L);
    rc = pull_state(b);
    if (rc) return rc;
    return check_err_bits(b->h_st->err);
}

int soil_hip_forward_comm_borders(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_set_rebuild_now(b->L);
    phase_forward(b->L);
    return 0;
}

int soil_hip_bin_neighbor_list(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    int rc = upload_geom_now(b);
    if (rc) return rc;
    phase_set_rebuild_now(b->L);
    for (int attempt = 0; attempt < 3; ++attempt) {
        phase_build(b->L, nullptr, nullptr, nullptr, nullptr);
        rc = pull_state(b);
        if (rc) return rc;
        if (!(b->h_st->err & ERR_PAIR_CAPACITY)) break;
        uint32_t np = b->h_st->npairs;
        rc = alloc_pairs(b, np + np / 2 + 4096, false);
        if (rc) return rc;
        b->h_st->err &= ~ERR_PAIR_CAPACITY;
        b->h_st->rebuilds -= 1;
        b->h_st->state = 0;
        HIPCHK(hipMemcpy(b->d_st, b->h_st, sizeof(BoxState), hipMemcpyHostToDevice));
        rc = push_desc(b);
        if (!rc) rc = rebuild_tiles(b);
        if (rc) return rc;
    }
    return check_err_bits(b->h_st->err);
}

int soil_hip_harmonic_force(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    if (!b->consumer) return set_error(SOIL_HIP_ESTATE, "consumer disabled (soil_hip_set_consumer)");
    phase_force(b->L);
    return 0;
}

int soil_hip_run_force_hook(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    if (!b->hook) return set_error(SOIL_HIP_ESTATE, "no force hook (soil_hip_set_force_hook)");
    int rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    return call_hook(b, true, b->h_st->steps);
}

int soil_hip_reverse_send_force(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    if (b->model.newton) {
        for (int d = 0; d < 3; ++d) b->L.rev_split[d] = 1;  // ordered rounds whatever the geometry
        phase_reverse(b->L);
    }
    return 0;
}

int soil_hip_final_integration(soil_hip_box* b) {
    PHASE_PROLOGUE(b);
    phase_final(b->L);
    return 0;
}

int soil_hip_step(soil_hip_box* b, int nsteps) {
    PHASE_PROLOGUE(b);
    if (nsteps < 0) return set_error(SOIL_HIP_EINVAL, "nsteps < 0");
    HIPCHK(hipSetDevice(b->ctx->device));
    return step_plain(b, nsteps);
}

int soil_hip_sync(soil_hip_box* b) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    int rc = pull_state(b);
    if (rc) return rc;
    if (b->timing) resolve_timing(b);
    return check_err_bits(b->h_st->err);
}

int soil_hip_set_tilt(soil_hip_box* b, double xy, double bvx) {
    PHASE_PROLOGUE(b);
    b->model.set_tilt(xy, bvx);
    return 0;
}

// ── readback ────────────────────────────────────────────────────────────────
int soil_hip_get_counts(soil_hip_box* b, uint64_t* out) {
    if (!b || !out) return set_error(SOIL_HIP_EINVAL, "null argument");
    if (!b->setup_done) {
        std::memset(out, 0, 8 * sizeof(uint64_t));
        out[0] = b->nlocal;
        return 0;
    }
    int rc = pull_state(b);
    if (rc) return rc;
    out[0] = b->h_st->nlocal;
    out[1] = b->h_st->nghost;
    out[2] = b->h_st->npairs;
    out[3] = b->model.ncells;
    out[4] = b->h_st->rebuilds;
    out[5] = b->h_st->steps;
    out[6] = b->model.stencil_fwd.size();
    out[7] = b->h_st->state;
    return 0;
}

int soil_hip_get_csr(soil_hip_box* b, uint32_t* offsets, size_t cap_off, uint32_t* indices, size_t cap_idx,
                     size_t* n_idx) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t nl = b->h_st->nlocal, np = b->h_st->npairs;
    if (n_idx) *n_idx = np;
    if (offsets) {
        if (cap_off < nl + 1) return set_error(SOIL_HIP_ECAPACITY, "offsets needs %zu", nl + 1);
        HIPCHK(hipMemcpy(offsets, b->h_dev.offsets, (nl + 1) * 4, hipMemcpyDeviceToHost));
    }
    if (indices) {
        if (cap_idx < np) return set_error(SOIL_HIP_ECAPACITY, "indices needs %zu", np);
        if (np) HIPCHK(hipMemcpy(indices, b->h_dev.indices, np * 4, hipMemcpyDeviceToHost));
    }
    return 0;
}

int soil_hip_classify_boundary_atoms(soil_hip_box* b, uint8_t* mask, size_t cap, size_t* n) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    const size_t nl = b->h_st->nlocal;
    if (n) *n = nl;
    if (!mask) return 0;
    if (cap < nl) return set_error(SOIL_HIP_ECAPACITY, "mask needs %zu", nl);
    if (b->h_st->rebuilds == 0) {  // no list built yet: every row is empty
        std::memset(mask, 0, nl);
        return 0;
    }
    phase_classify(b->L);
    rc = pull_desc(b);
    if (rc) return rc;
    std::vector<uint32_t> flags(nl);
    if (nl) HIPCHK(hipMemcpy(flags.data(), b->h_dev.slot, nl * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nl; ++i) mask[i] = (uint8_t)flags[i];
    return 0;
}

int soil_hip_box_device_bytes(soil_hip_box* b, size_t* estimate, size_t* allocated) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    if (estimate) {
        if (!b->setup_done && b->nlocal == 0) return set_error(SOIL_HIP_ESTATE, "no atoms uploaded");
        *estimate = b->setup_done ? b->est_bytes : box_bytes_estimate(b);
    }
    if (allocated) *allocated = box_bytes_allocated(b);
    return 0;
}

int soil_hip_compute_ke(soil_hip_box* b, const uint8_t* mask, double* out) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (!out) return set_error(SOIL_HIP_EINVAL, "null output");
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t nl = b->h_st->nlocal;
    // the box's own scratch: the mask bytes in the slot array, the partials in red
    uint8_t* dmask = mask ? (uint8_t*)b->h_dev.slot : nullptr;
    if (mask && nl) HIPCHK(hipMemcpy(dmask, mask, nl, hipMemcpyHostToDevice));
    phase_ke(b->L, dmask, (double*)b->h_dev.red);
    const size_t ntile = (nl + BLK - 1) / BLK;
    std::vector<double> part(ntile);
    HIPCHK(hipStreamSynchronize(b->stream));
    if (ntile) HIPCHK(hipMemcpy(part.data(), b->h_dev.red, ntile * 8, hipMemcpyDeviceToHost));
    double ke = 0.0;
    for (size_t k = 0; k < ntile; ++k) ke += part[k];
    *out = 0.5 * ke;
    return 0;
}

int soil_hip_compute_ke_noalloc(soil_hip_box* b, const uint8_t* mask, double* out) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (!out) return set_error(SOIL_HIP_EINVAL, "null output");
    rc = pull_state(b);  // nlocal (a copy into the box's pinned state)
    if (rc) return rc;
    const size_t nl = b->h_st->nlocal;
    // k_ke writes one partial per tile of the box's local capacity (nlocal <= it)
    if (b->caps.local > NOALLOC_MAX_LOCAL)
        return set_error(SOIL_HIP_ECAPACITY, "compute_ke_noalloc: %zu locals exceed the fixed workspace of %zu",
                         (size_t)b->caps.local, (size_t)NOALLOC_MAX_LOCAL);
    char* ws = (char*)scratch_ptr(b->ctx->device);
    if (!ws) return set_error(SOIL_HIP_EDEVICE, "compute_ke_noalloc: no workspace on device %d", b->ctx->device);
    uint8_t* dmask = mask ? (uint8_t*)(ws + NOALLOC_MASK_OFF) : nullptr;
    double* part = (double*)(ws + NOALLOC_PART_OFF);
    double* dsum = (double*)(ws + NOALLOC_OUT_OFF);
    if (mask && nl) HIPCHK(hipMemcpyAsync(dmask, mask, nl, hipMemcpyHostToDevice, b->stream));
    // every tile below nlocal writes its partial, and the sum writes the
    // result, so nothing left by an earlier call is read
    phase_ke(b->L, dmask, part);
    sum_tiles(part, (uint32_t)((nl + BLK - 1) / BLK), dsum, b->stream);
    HIPCHK(hipGetLastError());
    double ke = 0.0;
    HIPCHK(hipMemcpyAsync(&ke, dsum, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    *out = 0.5 * ke;
    return 0;
}

int soil_hip_get_atoms(soil_hip_box* b, size_t cap, double* pos, double* vel, double* force, double* radius,
                       double* mass, double* inv_mass, uint32_t* tag, int32_t* origin, int32_t* image) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t n = (size_t)b->h_st->nlocal + b->h_st->nghost;
    if (cap < n) return set_error(SOIL_HIP_ECAPACITY, "atoms needs %zu", n);
    const BoxDev& d = b->h_dev;
    std::vector<double> t0(n), t1(n), t2(n);
    auto get3d = [&](double* out, const double* a, const double* bb, const double* c) -> int {
        HIPCHK(hipMemcpy(t0.data(), a, n * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(t1.data(), bb, n * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(t2.data(), c, n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            out[3 * i] = t0[i];
            out[3 * i + 1] = t1[i];
            out[3 * i + 2] = t2[i];
        }
        return 0;
    };
    if (n == 0) return 0;
    if (pos && (rc = get3d(pos, d.x, d.y, d.z))) return rc;
    if (vel && (rc = get3d(vel, d.vx, d.vy, d.vz))) return rc;
    if (force && (rc = get3d(force, d.fx, d.fy, d.fz))) return rc;
    if (radius) HIPCHK(hipMemcpy(radius, d.rad, n * 8, hipMemcpyDeviceToHost));
    if (mass) HIPCHK(hipMemcpy(mass, d.mass, n * 8, hipMemcpyDeviceToHost));
    if (inv_mass) HIPCHK(hipMemcpy(inv_mass, d.inv_mass, n * 8, hipMemcpyDeviceToHost));
    if (tag) HIPCHK(hipMemcpy(tag, d.tag, n * 4, hipMemcpyDeviceToHost));
    if (origin) HIPCHK(hipMemcpy(origin, d.origin, n * 4, hipMemcpyDeviceToHost));
    if (image) {
        std::vector<int32_t> a(n), bb(n), c(n);
        HIPCHK(hipMemcpy(a.data(), d.imx, n * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(bb.data(), d.imy, n * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(c.data(), d.imz, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            image[3 * i] = a[i];
            image[3 * i + 1] = bb[i];
            image[3 * i + 2] = c[i];
        }
    }
    return 0;
}

int soil_hip_get_column(soil_hip_box* b, int col, double* out, size_t cap_atoms) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (col < 0 || col >= (int)b->cols.size()) return set_error(SOIL_HIP_EINVAL, "bad column");
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t n = (size_t)b->h_st->nlocal + b->h_st->nghost;
    if (cap_atoms < n) return set_error(SOIL_HIP_ECAPACITY, "column needs %zu atoms", n);
    if (n) HIPCHK(hipMemcpy(out, b->h_dev.cols[col].data, n * b->cols[col].width * 8, hipMemcpyDeviceToHost));
    return 0;
}

int soil_hip_get_types(soil_hip_box* b, uint32_t* type, size_t cap_atoms) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (!type) return set_error(SOIL_HIP_EINVAL, "null out");
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t n = (size_t)b->h_st->nlocal + b->h_st->nghost;
    if (cap_atoms < n) return set_error(SOIL_HIP_ECAPACITY, "types needs %zu atoms", n);
    if (n) HIPCHK(hipMemcpy(type, b->h_dev.type, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int soil_hip_get_bins(soil_hip_box* b, uint32_t* bin_start, size_t cap_cells1, uint32_t* sorted_atoms,
                      uint32_t* sorted_idx, uint32_t* atom_cell, size_t cap_atoms) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    rc = pull_desc(b);
    if (rc) return rc;
    const size_t n = (size_t)b->h_st->nlocal + b->h_st->nghost;
    const size_t c1 = (size_t)b->last_geom.ncells + 1;
    if (bin_start) {
        if (cap_cells1 < c1) return set_error(SOIL_HIP_ECAPACITY, "bin_start needs %zu", c1);
        HIPCHK(hipMemcpy(bin_start, b->h_dev.bin_start, c1 * 4, hipMemcpyDeviceToHost));
    }
    if ((sorted_atoms || sorted_idx || atom_cell) && cap_atoms < n)
        return set_error(SOIL_HIP_ECAPACITY, "atom arrays need %zu", n);
    if (n == 0) return 0;
    if (sorted_atoms) HIPCHK(hipMemcpy(sorted_atoms, b->h_dev.sorted, n * 4, hipMemcpyDeviceToHost));
    if (sorted_idx) HIPCHK(hipMemcpy(sorted_idx, b->h_dev.sidx, n * 4, hipMemcpyDeviceToHost));
    if (atom_cell) HIPCHK(hipMemcpy(atom_cell, b->h_dev.cell, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int soil_hip_get_swaps(soil_hip_box* b, double* out, size_t cap_rows, size_t* nrows) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    if (nrows) *nrows = SOIL_NSWAP;
    if (cap_rows < SOIL_NSWAP) return set_error(SOIL_HIP_ECAPACITY, "swaps need %d rows", SOIL_NSWAP);
    const Geom& g = b->last_geom;
    for (int s = 0; s < SOIL_NSWAP; ++s) {
        double* o = out + 11 * s;
        const int d = s / 2;
        o[0] = b->h_st->swap_start[s];
        o[1] = b->h_st->swap_count[s];
        o[2] = b->h_st->swap_send[s];
        o[3] = g.periodic[d] ? 0.0 : -1.0;
        o[4] = g.periodic[d] ? 0.0 : -1.0;
        for (int k = 0; k < 3; ++k) {
            o[5 + k] = g.swap_shift[s][k];
            o[8 + k] = g.swap_vshift[s][k];
        }
    }
    return 0;
}

int soil_hip_get_grid(soil_hip_box* b, double* out) {
    if (!b || !out) return set_error(SOIL_HIP_EINVAL, "null argument");
    const HostModel& m = b->model;
    for (int k = 0; k < 3; ++k) {
        out[k] = m.nbin[k];
        out[3 + k] = m.bin_size[k];
        out[6 + k] = m.bin_origin[k];
        out[12 + k] = m.tilt[k];
        out[15 + k] = m.bvel[k];
    }
    out[9] = m.ghost_cutoff;
    out[10] = m.cached_max_cutoff;
    out[11] = m.bin_min_size;
    out[18] = m.cached_min_skin;
    out[19] = m.uniform ? 1.0 : 0.0;
    out[20] = m.ucut_sq;
    return 0;
}

int soil_hip_get_stencil(soil_hip_box* b, int32_t* out, size_t cap, size_t* n) {
    if (!b) return set_error(SOIL_HIP_EINVAL, "null box");
    const auto& s = b->model.stencil_fwd;
    if (n) *n = s.size();
    if (out) {
        if (cap < s.size()) return set_error(SOIL_HIP_ECAPACITY, "stencil needs %zu", s.size());
        for (size_t i = 0; i < s.size(); ++i) out[i] = s[i];
    }
    return 0;
}

// Host-only: the grid and forward stencil a box with these radii gets at tilt xy.
int soil_hip_grid_for(const soil_hip_box_desc* desc, const double* radius, size_t n, double tilt_xy, double* out21,
                      int32_t* stencil, size_t cap, size_t* nstencil) {
    if (!desc || !out21 || (!radius && n)) return set_error(SOIL_HIP_EINVAL, "null argument");
    HostModel m;
    m.init(*desc);
    m.setup(radius, n);
    if (tilt_xy != 0.0 || desc->has_shear) {
        m.set_tilt(tilt_xy, 0.0);
        m.sort_prepare();
    }
    for (int k = 0; k < 3; ++k) {
        out21[k] = m.nbin[k];
        out21[3 + k] = m.bin_size[k];
        out21[6 + k] = m.bin_origin[k];
        out21[12 + k] = m.tilt[k];
        out21[15 + k] = m.bvel[k];
    }
    out21[9] = m.ghost_cutoff;
    out21[10] = m.cached_max_cutoff;
    out21[11] = m.bin_min_size;
    out21[18] = m.cached_min_skin;
    out21[19] = m.uniform ? 1.0 : 0.0;
    out21[20] = m.ucut_sq;
    const auto& s = m.newton ? m.stencil_fwd : m.stencil_full;
    if (nstencil) *nstencil = s.size();
    if (stencil) {
        if (cap < s.size()) return set_error(SOIL_HIP_ECAPACITY, "stencil needs %zu", s.size());
        for (size_t i = 0; i < s.size(); ++i) stencil[i] = s[i];
    }
    return 0;
}

// Copy kernel of known size (calibrates HBM traffic counters); returns ms per launch.
int soil_hip_probe_copy(soil_hip_ctx* ctx, size_t bytes, int reps, double* ms) {
    if (!ctx || bytes < 8 || reps < 1) return set_error(SOIL_HIP_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = bytes / 8;
    double *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 8));
    HIPCHK(hipMalloc(&b, n * 8));
    HIPCHK(hipMemset(a, 0, n * 8));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    copy_probe(a, b, n, nullptr);
    HIPCHK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) copy_probe(a, b, n, nullptr);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

// ── timing ─────────────────────────────────────────────────────────────────
int soil_hip_set_timing(soil_hip_box* b, int enabled) {
    int rc = check_box(b, true);
    if (rc) return rc;
    rc = pull_state(b);
    if (rc) return rc;
    resolve_timing(b);
    if (enabled < 0) return set_error(SOIL_HIP_EINVAL, "timing period %d < 0", enabled);
    b->timing = (uint32_t)enabled;
    b->timing_ctr = 0;
    b->t_builds = b->t_csr_ms = b->t_fill_ms = b->t_count_ms = 0.0;
    b->bytes_at_enable = b->h_st->csr_bytes;
    b->builds_at_enable = b->h_st->csr_builds;
    return 0;
}

int soil_hip_get_timing(soil_hip_box* b, double* out5) {
    int rc = check_box(b, true);
    if (rc) return rc;
    if (!out5) return set_error(SOIL_HIP_EINVAL, "null out");
    rc = pull_state(b);
    if (rc) return rc;
    resolve_timing(b);
    out5[0] = b->t_builds;
    out5[1] = b->t_csr_ms;
    // the timed builds' share of the algorithmic bytes of every build since
    // timing was set (at the average per build)
    const uint64_t nb = b->h_st->csr_builds - b->builds_at_enable;
    out5[2] = nb ? (double)(b->h_st->csr_bytes - b->bytes_at_enable) * (b->t_builds / (double)nb) : 0.0;
    out5[3] = b->t_fill_ms;
    out5[4] = b->t_count_ms;
    return 0;
}

}  // extern "C"
