// soil_ref — CPU restatement of SOIL's per-step particle substrate.
//
// TEST INFRASTRUCTURE ONLY. This is the oracle the HIP path is checked
// against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker or as the timed CPU baseline;
// the product library (soil_amd/) never calls it.
//
// Every function below restates one function of the reference (SueHeir/soil
// 0.1.3, Rust) literally, expression by expression, and cites the file:line
// it follows. All geometry is f64. The file must be compiled with
// -ffp-contract=off: the only fused multiply-adds are the explicit fma()
// calls that stand for Rust's `mul_add`.
//
// Parity pins (where each function is checked; see tests/):
//   x2lamda/lamda2x, pbc_triclinic  : domain.rs:1016-1058 KATs; every local
//                                     round-tripped through lamda, wrapped or not,
//                                     and x2lamda's left-to-right subtraction, on
//                                     atoms those misreadings would move
//                                     differently (domain.rs:211-229,686-712)
//   pbc / wrap_periodic             : the bound low + size computed at the wrap,
//                                     on atoms at hi and at lo + size where the
//                                     two differ (domain.rs:593-604)
//   initial/final_integration       : soil_verlet/src/lib.rs:221-233,416-428 KATs,
//                                     closed-form parabola, O(dt^2) ratio; an atom
//                                     whose first step a fused kick, a fused drift
//                                     or h = 0.5*dt/m would round differently
//   compute_bin_grid(_triclinic)    : closed-form grids, brute-force stencil
//                                     completeness (no in-cutoff pair missed);
//                                     cells by the product with 1/bin_size, on
//                                     atoms at faces where a division differs
//   bin_neighbor_list               : brute-force minimum-image pair sets (both
//                                     inclusions), neighbor.rs:1457-1619 fixtures;
//                                     row order: own cell first, then the stencil
//                                     in generation (dx, dy, dz loop) order, bin
//                                     order within a cell (neighbor.rs:1179-1304);
//                                     the uniform vs per-pair cutoff form and its
//                                     rounding, on a pair placed between the two
//                                     (neighbor.rs:679-688); the squared distance's
//                                     fma nesting, on pairs whose listing flips
//                                     under a plain sum or the reverse nesting
//   borders                         : brute-force image enumeration, swap rules;
//                                     face thresholds lo + gc / hi - gc rounded
//                                     once, on atoms where a difference form
//                                     would not send them (comm.rs:336-347); the
//                                     tilted low-x face (gc / Lx) * inflate, on an
//                                     atom gc * inflate / Lx would not send
//   forward_comm_borders            : every ghost = source + the build's fixed
//                                     offset, pos and vel, chained swap by swap
//                                     (comm.rs:400-432; tests/pins.py)
//   reverse_send_force + columns    : integer-valued #[reverse] ghost values sum
//                                     exactly onto their chain roots; #[forward]
//                                     ghost = root value; #[zero] cleared; a
//                                     tag-valued column follows its atom through
//                                     every sort (docs/SOIL_ATOMDATA_CONTRACT.md
//                                     :87-143); force balance is exact; with
//                                     inexact values, the sums in the reference's
//                                     order (rounds last to first, low swap then
//                                     high, comm.rs:962-1021) on a box whose z
//                                     selections overlap and whose ghosts chain
//   sort_atoms_by_bin               : stable-sort invariants vs numpy argsort; a
//                                     periodic sort rebuilds exactly when that
//                                     order is not the identity (neighbor.rs:989-996)
//   decide_rebuild / every / check  : every=E, check off rebuilds on steps 1,
//                                     E+2, 2E+3, ... (neighbor.rs:762-784); the
//                                     threshold fires at ((sf-1) min_r)^2; the
//                                     squared displacement fused when the box is
//                                     fully periodic, a plain sum otherwise, on
//                                     atoms whose decision the two forms split
//   apply_deform (LE xy)            : closed form soil_deform/src/lib.rs:576-583;
//                                     the left-to-right product of :579, on a rate
//                                     and step every regrouping rounds differently;
//                                     the streaming profile rate * (y - yc) (:568-574)
//   step state machine              : the run_if placement is read from the
//                                     plugin sources (tests/test_integration_doc.py
//                                     parses them); the grass scheduler's dispatch
//                                     timing is external to the reference tree. A
//                                     periodic sort that rebuilds a CommunicateOnly
//                                     step leaves an atom past a face unwrapped
//                                     (pbc's guard was already evaluated)

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <cfloat>

// memcpy that accepts an empty range (an empty vector's data() may be null)
static inline void copy_bytes(void* dst, const void* src, size_t bytes) {
    if (bytes) std::memcpy(dst, src, bytes);
}

namespace {

enum CommState { FullRebuild = 0, CommunicateOnly = 1 };

// AtomData extension column (soil_derive contract, docs/SOIL_ATOMDATA_CONTRACT.md:87-143):
// width f64 per atom; FORWARD = overwrite on ghosts, REVERSE = += onto owner,
// ZERO = zeroed every step.
enum { COL_FORWARD = 1, COL_REVERSE = 2, COL_ZERO = 4 };
struct Column {
    int width = 1;
    unsigned flags = 0;
    std::vector<double> data;  // len = width * natoms
};

// atom.rs:406-431 (SoA of per-atom fields; Real = Accum = f64)
struct Atom {
    uint32_t nlocal = 0, nghost = 0;
    double dt = 1.0;
    std::vector<uint32_t> tag, atom_type;
    std::vector<int32_t> origin_index;
    std::vector<uint8_t> is_ghost;
    std::vector<double> pos, vel, force;  // interleaved [x,y,z]
    std::vector<double> cutoff_radius, mass, inv_mass;
    std::vector<int32_t> image;           // interleaved
    size_t len() const { return cutoff_radius.size(); }
};

// domain.rs:105-142
struct Domain {
    double boundaries_low[3] = {0, 0, 0}, boundaries_high[3] = {1, 1, 1};
    double sub_domain_low[3] = {0, 0, 0}, sub_domain_high[3] = {1, 1, 1};
    double sub_length[3] = {1, 1, 1}, size[3] = {1, 1, 1};
    int periodic[3] = {1, 1, 1};
    bool shrink[3] = {false, false, false};  // BoundaryType::ShrinkWrap (domain.rs:27-34)
    double shrink_wrap_padding = 0.0;
    bool bounds_changed = false;
    double ghost_cutoff = 0.0;
    double tilt[3] = {0, 0, 0};
    bool triclinic = false;
    double boundary_vel[3] = {0, 0, 0};
    double sub_lamda_low[3] = {0, 0, 0}, sub_lamda_high[3] = {1, 1, 1};

    // domain.rs:211-217
    void x2lamda(const double r[3], double out[3]) const {
        double xy = tilt[0], xz = tilt[1], yz = tilt[2];
        double lz = (r[2] - boundaries_low[2]) / size[2];
        double ly = (r[1] - boundaries_low[1] - yz * lz) / size[1];
        double lx = (r[0] - boundaries_low[0] - xy * ly - xz * lz) / size[0];
        out[0] = lx; out[1] = ly; out[2] = lz;
    }
    // domain.rs:222-229
    void lamda2x(const double lam[3], double out[3]) const {
        double xy = tilt[0], xz = tilt[1], yz = tilt[2];
        out[0] = boundaries_low[0] + size[0] * lam[0] + xy * lam[1] + xz * lam[2];
        out[1] = boundaries_low[1] + size[1] * lam[1] + yz * lam[2];
        out[2] = boundaries_low[2] + size[2] * lam[2];
    }
};

// neighbor.rs:87-174 (only the fields the path reads)
struct Neighbor {
    double skin_fraction = 1.0;
    std::vector<uint32_t> neighbor_offsets, neighbor_indices;
    double bin_min_size = 1.0;
    double bin_size[3] = {1, 1, 1};
    int bin_count[3] = {1, 1, 1};
    std::vector<double> last_build_pos;  // interleaved, len = 3*n
    size_t steps_since_build = 0;
    size_t last_build_total = 0;
    double pbc_box[3] = {0, 0, 0};
    bool pbc_flags[3] = {false, false, false};
    std::vector<int32_t> bin_stencil, bin_stencil_forward;
    bool bin_stencil_self = false;
    double bin_origin[3] = {0, 0, 0};
    size_t bin_total_cells = 0;
    size_t every = 0;
    bool check = true;
    double ghost_cutoff = 0.0;
    double cached_min_skin = DBL_MAX;
    size_t sort_every = 1000;
    size_t sort_counter = 0;
    std::vector<uint32_t> bin_atom_cell, bin_start, bin_sorted_atoms, bin_atom_sorted_idx;
    std::vector<double> bin_sorted_pos;
    std::vector<uint32_t> sorted_local_cell;
    bool local_cells_valid = false;
    bool has_uniform_cutoff = false;
    double cached_uniform_cutoff_sq = 0.0;
    double cached_max_cutoff = 0.0;
    bool newton = true;
    // observability of the last sort (not reference state)
    std::vector<uint64_t> last_sort_perm;
    bool last_sort_ran = false;
};

// comm.rs:114-122
struct SwapData {
    std::vector<size_t> send_indices;
    size_t recv_start = 0, recv_count = 0;
    int to_proc = -1, from_proc = -1;
    double periodic_offset[3] = {0, 0, 0};
    double vel_offset[3] = {0, 0, 0};
};

// comm.rs:128-133
struct CommTopology {
    int swap_directions[2][3] = {{-1, -1, -1}, {-1, -1, -1}};
    double periodic_swap[2][3] = {{0, 0, 0}, {0, 0, 0}};
    std::vector<SwapData> swap_data;
    int maxneed[3] = {0, 0, 0};
};

// soil_deform/src/lib.rs:195-211 (xy Lees–Edwards part)
// soil_deform/src/lib.rs:168-205: per-axis style (-1 none, 0 erate, 1 vel,
// 2 final) with its parameter(s) and the bounds captured at setup
struct AxisDeform {
    int style = -1;
    double value = 0.0;  // erate: rate, vel: velocity
    double target_lo = 0.0, target_hi = 0.0;
    double lo_0 = 0.0, hi_0 = 0.0;
};

struct DeformState {
    AxisDeform axes[3];
    bool has_shear = false;
    double shear_xy_rate = 0.0;
    bool shear_initialized = false;
    bool remap = true;
    size_t step = 0;
    bool initialized = false;
    size_t stage_steps = 1;  // cycle_count + cycle_remaining of the stage (the "final" style's horizon)
    bool has_any() const {
        return has_shear || axes[0].style >= 0 || axes[1].style >= 0 || axes[2].style >= 0;
    }
};

}  // namespace

extern "C" {

// Box description shared with the tests (field order is the C layout).
typedef struct {
    double lo[3], hi[3];
    int32_t periodic[3];
    int32_t newton;
    double skin_fraction;
    double bin_min_size;
    uint64_t every;
    int32_t check;
    int32_t pad0;
    uint64_t sort_every;
    double dt;
    double shear_rate;      // Lees–Edwards xy erate; used when has_shear
    int32_t has_shear;
    int32_t pad1;
} soil_ref_desc;

struct soil_ref_box {
    Atom atoms;
    Domain domain;
    Neighbor neighbor;
    CommTopology topo;
    DeformState deform;
    CommState state = FullRebuild;
    std::vector<Column> cols;
    double harmonic_k = 0.0;
    bool use_consumer = false;
    bool collapsed = false;  // a deformed axis reached zero size (the reference panics)
    uint64_t rebuilds = 0;
    uint64_t steps = 0;
};

}  // extern "C"

namespace {

const int RANK = 0;  // single process

// ── column (registry) helpers: atom.rs:163-377 ──────────────────────────────
void cols_truncate(soil_ref_box* b, size_t n) {
    for (auto& c : b->cols) c.data.resize(n * c.width);
}
void cols_push_copy(soil_ref_box* b, size_t src) {
    // pack_all + unpack_all of a ghost: copies every field of atom src
    for (auto& c : b->cols)
        for (int w = 0; w < c.width; ++w) c.data.push_back(c.data[src * c.width + w]);
}

void atoms_truncate_to_nlocal(Atom& a) {
    size_t n = a.nlocal;
    a.tag.resize(n); a.atom_type.resize(n); a.origin_index.resize(n); a.is_ghost.resize(n);
    a.pos.resize(3 * n); a.vel.resize(3 * n); a.force.resize(3 * n);
    a.cutoff_radius.resize(n); a.mass.resize(n); a.inv_mass.resize(n); a.image.resize(3 * n);
}

// atom.rs:625-629
void remove_ghost_atoms(soil_ref_box* b) {
    atoms_truncate_to_nlocal(b->atoms);
    cols_truncate(b, b->atoms.nlocal);
    b->atoms.nghost = 0;
}

// atom.rs:631-635
void zero_all_forces(soil_ref_box* b) {
    size_t n = b->atoms.len();
    for (size_t i = 0; i < 3 * n; ++i) b->atoms.force[i] = 0.0;
    for (auto& c : b->cols)
        if (c.flags & COL_ZERO) {
            c.data.resize(n * c.width);
            std::fill(c.data.begin(), c.data.end(), 0.0);
        }
}

// neighbor.rs:311-397
void compute_bin_grid_triclinic(Neighbor& nb, const Domain& d, double max_cutoff, double min_bin);
void compute_bin_grid(Neighbor& nb, const Domain& d, int comm_size) {
    double max_cutoff = nb.cached_max_cutoff;
    double required_bin = comm_size > 1 ? max_cutoff * 0.5 : max_cutoff;
    double min_bin = std::max(nb.bin_min_size, required_bin);
    if (d.triclinic) {
        compute_bin_grid_triclinic(nb, d, max_cutoff, min_bin);
        return;
    }
    int xi = (int)std::max(std::floor(d.sub_length[0] / min_bin), 1.0);
    int yi = (int)std::max(std::floor(d.sub_length[1] / min_bin), 1.0);
    int zi = (int)std::max(std::floor(d.sub_length[2] / min_bin), 1.0);
    double abs_[3] = {d.sub_length[0] / (double)xi, d.sub_length[1] / (double)yi,
                      d.sub_length[2] / (double)zi};
    double ghost_cut = d.ghost_cutoff > max_cutoff ? d.ghost_cutoff : max_cutoff;
    int srx = (int)std::ceil(max_cutoff / abs_[0]);
    int sry = (int)std::ceil(max_cutoff / abs_[1]);
    int srz = (int)std::ceil(max_cutoff / abs_[2]);
    int sx = (int)std::ceil(ghost_cut / abs_[0]) + srx;
    int sy = (int)std::ceil(ghost_cut / abs_[1]) + sry;
    int sz = (int)std::ceil(ghost_cut / abs_[2]) + srz;
    int nx = xi + 2 * sx, ny = yi + 2 * sy, nz = zi + 2 * sz;
    nb.bin_count[0] = nx; nb.bin_count[1] = ny; nb.bin_count[2] = nz;
    for (int k = 0; k < 3; ++k) nb.bin_size[k] = abs_[k];
    nb.bin_total_cells = (size_t)(nx * ny * nz);
    nb.bin_origin[0] = d.sub_domain_low[0] - abs_[0] * (double)sx;
    nb.bin_origin[1] = d.sub_domain_low[1] - abs_[1] * (double)sy;
    nb.bin_origin[2] = d.sub_domain_low[2] - abs_[2] * (double)sz;
    double cutoff_sq = max_cutoff * max_cutoff;
    nb.bin_stencil.clear();
    nb.bin_stencil_forward.clear();
    nb.bin_stencil_self = false;
    for (int dx = -sx; dx <= sx; ++dx)
        for (int dy = -sy; dy <= sy; ++dy)
            for (int dz = -sz; dz <= sz; ++dz) {
                double mdx = (double)std::max(std::abs(dx) - 1, 0) * abs_[0];
                double mdy = (double)std::max(std::abs(dy) - 1, 0) * abs_[1];
                double mdz = (double)std::max(std::abs(dz) - 1, 0) * abs_[2];
                if (mdx * mdx + mdy * mdy + mdz * mdz < cutoff_sq) {
                    int off = dx * ny * nz + dy * nz + dz;
                    nb.bin_stencil.push_back(off);
                    if (off > 0) nb.bin_stencil_forward.push_back(off);
                    else if (off == 0) nb.bin_stencil_self = true;
                }
            }
    nb.pbc_box[0] = d.boundaries_high[0] - d.boundaries_low[0];
    nb.pbc_box[1] = d.boundaries_high[1] - d.boundaries_low[1];
    nb.pbc_box[2] = d.boundaries_high[2] - d.boundaries_low[2];
}

// neighbor.rs:411-495
void compute_bin_grid_triclinic(Neighbor& nb, const Domain& d, double max_cutoff, double min_bin) {
    const double* l = d.size;
    double xy = d.tilt[0];
    double sub_l[3] = {d.sub_lamda_high[0] - d.sub_lamda_low[0], d.sub_lamda_high[1] - d.sub_lamda_low[1],
                       d.sub_lamda_high[2] - d.sub_lamda_low[2]};
    int ni[3];
    for (int k = 0; k < 3; ++k) ni[k] = (int)std::max(std::floor(sub_l[k] * l[k] / min_bin), 1.0);
    double bl[3] = {sub_l[0] / (double)ni[0], sub_l[1] / (double)ni[1], sub_l[2] / (double)ni[2]};
    double real_bin[3] = {bl[0] * l[0], bl[1] * l[1], bl[2] * l[2]};
    double ghost_cut = d.ghost_cutoff > max_cutoff ? d.ghost_cutoff : max_cutoff;
    int sr[3];
    for (int k = 0; k < 3; ++k) sr[k] = (int)std::ceil(max_cutoff / real_bin[k]);
    int skew_x_bins = (int)std::ceil(std::fabs(xy) / real_bin[0]);
    int sx = (int)std::ceil(ghost_cut / real_bin[0]) + sr[0] + skew_x_bins;
    int sy = (int)std::ceil(ghost_cut / real_bin[1]) + sr[1];
    int sz = (int)std::ceil(ghost_cut / real_bin[2]) + sr[2];
    int nx = ni[0] + 2 * sx, ny = ni[1] + 2 * sy, nz = ni[2] + 2 * sz;
    nb.bin_count[0] = nx; nb.bin_count[1] = ny; nb.bin_count[2] = nz;
    for (int k = 0; k < 3; ++k) nb.bin_size[k] = bl[k];
    nb.bin_total_cells = (size_t)(nx * ny * nz);
    nb.bin_origin[0] = d.sub_lamda_low[0] - bl[0] * (double)sx;
    nb.bin_origin[1] = d.sub_lamda_low[1] - bl[1] * (double)sy;
    nb.bin_origin[2] = d.sub_lamda_low[2] - bl[2] * (double)sz;
    double cutoff_sq = max_cutoff * max_cutoff;
    nb.bin_stencil.clear();
    nb.bin_stencil_forward.clear();
    nb.bin_stencil_self = false;
    for (int dx = -sx; dx <= sx; ++dx)
        for (int dy = -sy; dy <= sy; ++dy)
            for (int dz = -sz; dz <= sz; ++dz) {
                double m1 = (double)std::max(std::abs(dy) - 1, 0) * bl[1];
                double m2 = (double)std::max(std::abs(dz) - 1, 0) * bl[2];
                double dyz_sq = (l[1] * m1) * (l[1] * m1) + (l[2] * m2) * (l[2] * m2);
                if (dyz_sq >= cutoff_sq) continue;
                double budget = std::sqrt(cutoff_sq - dyz_sq);
                double a_lo = l[0] * ((double)dx - 1.0) * bl[0];
                double a_hi = l[0] * ((double)dx + 1.0) * bl[0];
                double b1 = xy * ((double)dy - 1.0) * bl[1];
                double b2 = xy * ((double)dy + 1.0) * bl[1];
                double b_lo = b1 <= b2 ? b1 : b2;
                double b_hi = b1 <= b2 ? b2 : b1;
                double lo = a_lo + b_lo;
                double hi = a_hi + b_hi;
                if (hi >= -budget && lo <= budget) {
                    int off = dx * ny * nz + dy * nz + dz;
                    nb.bin_stencil.push_back(off);
                    if (off > 0) nb.bin_stencil_forward.push_back(off);
                    else if (off == 0) nb.bin_stencil_self = true;
                }
            }
    nb.pbc_box[0] = l[0]; nb.pbc_box[1] = l[1]; nb.pbc_box[2] = l[2];
}

// neighbor.rs:498-503
void recompute_bins(Neighbor& nb, const Domain& d, int comm_size) {
    if (nb.cached_max_cutoff <= 0.0) return;
    compute_bin_grid(nb, d, comm_size);
}

// neighbor.rs:604-665
void neighbor_setup(soil_ref_box* b) {
    Neighbor& nb = b->neighbor;
    Domain& d = b->domain;
    Atom& a = b->atoms;
    double local_max_skin = 0.0;
    for (double c : a.cutoff_radius) local_max_skin = std::fmax(local_max_skin, c);
    double max_skin = local_max_skin;  // -all_reduce_min(-x) on 1 rank
    double max_cutoff = max_skin > 0.0 ? 2.0 * max_skin * nb.skin_fraction : nb.bin_min_size;
    double local_min_skin = DBL_MAX;
    for (double c : a.cutoff_radius) local_min_skin = std::fmin(local_min_skin, c);
    double min_skin = local_min_skin;
    double displacement_buffer = min_skin < DBL_MAX * 0.5 ? (nb.skin_fraction - 1.0) * min_skin : 0.0;
    double ghost_cut = std::fmax(max_cutoff + 2.0 * displacement_buffer, d.ghost_cutoff);
    nb.ghost_cutoff = ghost_cut;
    nb.cached_max_cutoff = max_cutoff;
    d.ghost_cutoff = ghost_cut;
    double required_bin = max_cutoff;  // comm.size() == 1
    if (nb.bin_min_size < required_bin) nb.bin_min_size = required_bin;
    compute_bin_grid(nb, d, 1);
    for (int k = 0; k < 3; ++k) nb.pbc_flags[k] = d.periodic[k] != 0;
}

// comm.rs:253-283 (single rank: rank 0 in a 1x1x1 grid)
void comm_setup(soil_ref_box* b) {
    for (int dim = 0; dim < 3; ++dim) {
        // forward neighbour: pos+1 < 1 is false
        if (b->domain.periodic[dim]) {
            b->topo.swap_directions[1][dim] = RANK;
            b->topo.periodic_swap[1][dim] = -1.0;
        }
        if (b->domain.periodic[dim]) {
            b->topo.swap_directions[0][dim] = RANK;
            b->topo.periodic_swap[0][dim] = 1.0;
        }
    }
}

// neighbor.rs:669-690
void save_build_positions(soil_ref_box* b) {
    Atom& a = b->atoms;
    Neighbor& nb = b->neighbor;
    size_t nlocal = a.nlocal;
    nb.last_build_pos.assign(a.pos.begin(), a.pos.begin() + 3 * nlocal);
    nb.last_build_total = a.len();
    nb.steps_since_build = 0;
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (size_t i = 0; i < nlocal; ++i) {
        mn = std::fmin(mn, a.cutoff_radius[i]);
        mx = std::fmax(mx, a.cutoff_radius[i]);
    }
    nb.cached_min_skin = mn;
    if (std::fabs(mx - mn) < 1e-15) {
        double cutoff = 2.0 * mn * nb.skin_fraction;
        nb.has_uniform_cutoff = true;
        nb.cached_uniform_cutoff_sq = cutoff * cutoff;
    } else {
        nb.has_uniform_cutoff = false;
    }
}

// neighbor.rs:693-758
bool displacement_exceeded(const soil_ref_box* b) {
    const Atom& a = b->atoms;
    const Neighbor& nb = b->neighbor;
    double min_r = nb.cached_min_skin;
    double per_atom_skin = (nb.skin_fraction - 1.0) * min_r;
    double threshold_sq = per_atom_skin * per_atom_skin;
    double bx = nb.pbc_box[0], by = nb.pbc_box[1], bz = nb.pbc_box[2];
    bool px = nb.pbc_flags[0], py = nb.pbc_flags[1], pz = nb.pbc_flags[2];
    size_t n = nb.last_build_pos.size() / 3;
    if (px && py && pz) {
        double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
        for (size_t i = 0; i < n; ++i) {
            double dx = a.pos[3 * i] - nb.last_build_pos[3 * i];
            double dy = a.pos[3 * i + 1] - nb.last_build_pos[3 * i + 1];
            double dz = a.pos[3 * i + 2] - nb.last_build_pos[3 * i + 2];
            dx -= bx * std::round(dx * ibx);
            dy -= by * std::round(dy * iby);
            dz -= bz * std::round(dz * ibz);
            if (std::fma(dx, dx, std::fma(dy, dy, dz * dz)) > threshold_sq) return true;
        }
        return false;
    }
    double hbx = bx * 0.5, hby = by * 0.5, hbz = bz * 0.5;
    for (size_t i = 0; i < n; ++i) {
        double dx = a.pos[3 * i] - nb.last_build_pos[3 * i];
        double dy = a.pos[3 * i + 1] - nb.last_build_pos[3 * i + 1];
        double dz = a.pos[3 * i + 2] - nb.last_build_pos[3 * i + 2];
        if (px) { if (dx > hbx) dx -= bx; else if (dx < -hbx) dx += bx; }
        if (py) { if (dy > hby) dy -= by; else if (dy < -hby) dy += by; }
        if (pz) { if (dz > hbz) dz -= bz; else if (dz < -hbz) dz += bz; }
        if (dx * dx + dy * dy + dz * dz > threshold_sq) return true;
    }
    return false;
}

// neighbor.rs:762-784
bool needs_rebuild(const soil_ref_box* b) {
    size_t nlocal = b->atoms.nlocal;
    const Neighbor& nb = b->neighbor;
    if (nb.last_build_pos.size() / 3 != nlocal || nlocal == 0) return true;
    if (b->state == FullRebuild) return true;
    if (nb.every == 0) return displacement_exceeded(b);
    if (nb.check) return nb.steps_since_build >= nb.every || displacement_exceeded(b);
    return nb.steps_since_build >= nb.every;
}

// neighbor.rs:796-822
void decide_rebuild(soil_ref_box* b) {
    bool local_needs = needs_rebuild(b);
    if (local_needs) b->state = FullRebuild;
    else b->neighbor.steps_since_build += 1;
}

// cell index of a position (neighbor.rs:915-922 and 1065-1076). Rust `as i32`
// saturates and maps NaN to 0; the clamp to [0, n-1] follows.
inline int cell_coord(double v, int n) {
    double f = std::floor(v);
    if (!(f >= 0.0)) return 0;            // negative or NaN
    if (f >= (double)(n - 1)) return n - 1;
    return (int)f;
}

// atom.rs:479-490 + atom.rs:291-295
void apply_permutation(soil_ref_box* b, const std::vector<size_t>& perm, size_t n) {
    Atom& a = b->atoms;
    auto perm1 = [&](auto& v) {
        auto scratch = v;
        for (size_t k = 0; k < n; ++k) v[k] = scratch[perm[k]];
    };
    auto perm3 = [&](auto& v) {
        auto scratch = v;
        for (size_t k = 0; k < n; ++k)
            for (int c = 0; c < 3; ++c) v[3 * k + c] = scratch[3 * perm[k] + c];
    };
    perm1(a.tag); perm1(a.atom_type); perm1(a.origin_index); perm1(a.is_ghost);
    perm3(a.pos); perm3(a.vel); perm3(a.force);
    perm1(a.cutoff_radius); perm1(a.mass); perm1(a.inv_mass); perm3(a.image);
    for (auto& c : b->cols) {
        auto scratch = c.data;
        for (size_t k = 0; k < n; ++k)
            for (int w = 0; w < c.width; ++w) c.data[k * c.width + w] = scratch[perm[k] * c.width + w];
    }
}

// neighbor.rs:834-997
void sort_atoms_by_bin(soil_ref_box* b) {
    Neighbor& nb = b->neighbor;
    Domain& d = b->domain;
    Atom& a = b->atoms;
    nb.last_sort_ran = false;
    if (d.bounds_changed) {
        recompute_bins(nb, d, 1);
        d.bounds_changed = false;
        b->state = FullRebuild;
    }
    nb.sort_counter += 1;
    bool periodic_sort = nb.sort_every > 0 && nb.sort_counter >= nb.sort_every;
    if (periodic_sort) nb.sort_counter = 0;
    size_t nlocal = a.nlocal;
    nb.local_cells_valid = false;
    if (nlocal == 0 || nb.bin_total_cells == 0 || nlocal > a.len()) {
        // global_did_sort == 0 on one rank: nothing to do
        return;
    }
    bool full_rebuild = b->state == FullRebuild;
    if (!periodic_sort && !full_rebuild) return;

    double inv_bsx = 1.0 / nb.bin_size[0], inv_bsy = 1.0 / nb.bin_size[1], inv_bsz = 1.0 / nb.bin_size[2];
    int nx = nb.bin_count[0], ny = nb.bin_count[1], nz = nb.bin_count[2];
    size_t total_cells = (size_t)(nx * ny * nz);
    double ox = nb.bin_origin[0], oy = nb.bin_origin[1], oz = nb.bin_origin[2];
    std::vector<uint32_t> cell(nlocal, 0), counts(total_cells + 1, 0);
    std::vector<size_t> perm(nlocal, 0);
    bool triclinic = d.triclinic;
    for (size_t i = 0; i < nlocal; ++i) {
        double p[3] = {a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2]};
        if (triclinic) { double q[3]; d.x2lamda(p, q); p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; }
        int cx = cell_coord((p[0] - ox) * inv_bsx, nx);
        int cy = cell_coord((p[1] - oy) * inv_bsy, ny);
        int cz = cell_coord((p[2] - oz) * inv_bsz, nz);
        uint32_t c = (uint32_t)(cx * ny * nz + cy * nz + cz);
        cell[i] = c;
        counts[c + 1] += 1;
    }
    for (size_t c = 0; c < total_cells; ++c) counts[c + 1] += counts[c];
    for (size_t i = 0; i < nlocal; ++i) {
        uint32_t c = cell[i];
        uint32_t slot = counts[c];
        perm[slot] = i;
        counts[c] = slot + 1;
    }
    bool already_sorted = true;
    for (size_t i = 0; i < nlocal; ++i)
        if (perm[i] != i) { already_sorted = false; break; }
    nb.sorted_local_cell.assign(nlocal, 0);
    for (size_t k = 0; k < nlocal; ++k) nb.sorted_local_cell[k] = cell[perm[k]];
    nb.local_cells_valid = true;
    nb.last_sort_perm.assign(perm.begin(), perm.end());
    nb.last_sort_ran = true;
    if (!already_sorted) {
        apply_permutation(b, perm, nlocal);
        if (nb.last_build_pos.size() / 3 >= nlocal) {
            std::vector<double> scratch(nb.last_build_pos.begin(), nb.last_build_pos.begin() + 3 * nlocal);
            for (size_t k = 0; k < nlocal; ++k)
                for (int c = 0; c < 3; ++c) nb.last_build_pos[3 * k + c] = scratch[3 * perm[k] + c];
        }
    }
    if (periodic_sort) {
        if (!already_sorted) {
            nb.last_build_pos.clear();
            b->state = FullRebuild;
        }
    }
}

// neighbor.rs:1013-1319 (newton = true half list and newton = false full list)
void bin_neighbor_list(soil_ref_box* b) {
    Atom& a = b->atoms;
    Neighbor& nb = b->neighbor;
    const Domain& d = b->domain;
    size_t nlocal = a.nlocal;
    size_t total = a.len();
    bool triclinic = d.triclinic;
    b->state = CommunicateOnly;
    save_build_positions(b);
    int ny = nb.bin_count[1], nz = nb.bin_count[2], nx = nb.bin_count[0];
    size_t total_cells = nb.bin_total_cells;
    double inv_bsx = 1.0 / nb.bin_size[0], inv_bsy = 1.0 / nb.bin_size[1], inv_bsz = 1.0 / nb.bin_size[2];
    double ox = nb.bin_origin[0], oy = nb.bin_origin[1], oz = nb.bin_origin[2];
    std::vector<uint32_t>& atom_cell = nb.bin_atom_cell;
    atom_cell.assign(total, 0);
    std::vector<uint32_t> bin_count_arr(total_cells, 0);
    bool reuse_local = nb.local_cells_valid && nb.sorted_local_cell.size() == nlocal;
    size_t ghost_start = 0;
    if (reuse_local) {
        for (size_t i = 0; i < nlocal; ++i) {
            uint32_t c = nb.sorted_local_cell[i];
            atom_cell[i] = c;
            bin_count_arr[c] += 1;
        }
        ghost_start = nlocal;
    }
    for (size_t i = ghost_start; i < total; ++i) {
        double p[3] = {a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2]};
        if (triclinic) { double q[3]; d.x2lamda(p, q); p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; }
        int cx = cell_coord((p[0] - ox) * inv_bsx, nx);
        int cy = cell_coord((p[1] - oy) * inv_bsy, ny);
        int cz = cell_coord((p[2] - oz) * inv_bsz, nz);
        uint32_t c = (uint32_t)(cx * ny * nz + cy * nz + cz);
        atom_cell[i] = c;
        bin_count_arr[c] += 1;
    }
    nb.local_cells_valid = false;
    std::vector<uint32_t>& bin_start = nb.bin_start;
    bin_start.assign(total_cells + 1, 0);
    for (size_t c = 0; c < total_cells; ++c) bin_start[c + 1] = bin_start[c] + bin_count_arr[c];
    std::vector<uint32_t>& sorted_atoms = nb.bin_sorted_atoms;
    std::vector<uint32_t>& atom_sorted_idx = nb.bin_atom_sorted_idx;
    sorted_atoms.assign(total, 0);
    atom_sorted_idx.assign(total, 0);
    for (size_t c = 0; c < total_cells; ++c) bin_count_arr[c] = bin_start[c];
    for (size_t i = 0; i < total; ++i) {
        uint32_t c = atom_cell[i];
        uint32_t p = bin_count_arr[c];
        sorted_atoms[p] = (uint32_t)i;
        atom_sorted_idx[i] = p;
        bin_count_arr[c] = p + 1;
    }
    std::vector<double>& sorted_pos = nb.bin_sorted_pos;
    sorted_pos.assign(3 * total, 0.0);
    for (size_t m = 0; m < total; ++m)
        for (int c = 0; c < 3; ++c) sorted_pos[3 * m + c] = a.pos[3 * sorted_atoms[m] + c];

    bool has_self = nb.bin_stencil_self;
    double skin_fraction = nb.skin_fraction;
    bool newton = nb.newton;
    const std::vector<int32_t>& stencil = newton ? nb.bin_stencil_forward : nb.bin_stencil;
    std::vector<uint32_t>& offsets = nb.neighbor_offsets;
    std::vector<uint32_t>& indices = nb.neighbor_indices;
    offsets.clear();
    indices.clear();
    offsets.reserve(nlocal + 1);
    bool uniform = nb.has_uniform_cutoff;
    double cutoff_sq = nb.cached_uniform_cutoff_sq;
    double skin_fraction_sq = skin_fraction * skin_fraction;
    for (size_t i = 0; i < nlocal; ++i) {
        offsets.push_back((uint32_t)indices.size());
        size_t my_cell = atom_cell[i];
        double pi0 = a.pos[3 * i], pi1 = a.pos[3 * i + 1], pi2 = a.pos[3 * i + 2];
        double si = a.cutoff_radius[i];
        auto accept = [&](size_t m, size_t j) -> bool {
            double dx = sorted_pos[3 * m] - pi0;
            double dy = sorted_pos[3 * m + 1] - pi1;
            double dz = sorted_pos[3 * m + 2] - pi2;
            double r2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
            if (uniform) return r2 < cutoff_sq;
            double sum_skin = si + a.cutoff_radius[j];
            return r2 < sum_skin * sum_skin * skin_fraction_sq;
        };
        if (has_self) {
            size_t cell_start = bin_start[my_cell];
            size_t end = bin_start[my_cell + 1];
            if (newton) {
                size_t self_start = (size_t)atom_sorted_idx[i] + 1;
                for (size_t m = self_start; m < end; ++m) {
                    size_t j = sorted_atoms[m];
                    if (accept(m, j)) indices.push_back((uint32_t)j);
                }
            } else {
                for (size_t m = cell_start; m < end; ++m) {
                    size_t j = sorted_atoms[m];
                    if (j == i) continue;
                    if (accept(m, j)) indices.push_back((uint32_t)j);
                }
            }
        }
        for (int32_t off : stencil) {
            if (!newton && off == 0) continue;
            size_t c = (size_t)((int64_t)my_cell + off);
            size_t start = bin_start[c], end = bin_start[c + 1];
            for (size_t m = start; m < end; ++m) {
                size_t j = sorted_atoms[m];
                if (accept(m, j)) indices.push_back((uint32_t)j);
            }
        }
    }
    offsets.push_back((uint32_t)indices.size());
    b->rebuilds += 1;
}

// domain.rs:594-605
inline double wrap_periodic(double pos, double low, double size, int* delta) {
    double high = low + size;
    if (pos < low) { pos += size; *delta = -1; }
    else if (pos >= high) { pos -= size; *delta = 1; }
    else *delta = 0;
    return pos;
}

void atom_swap_remove(soil_ref_box* b, size_t i) {
    Atom& a = b->atoms;
    size_t last = a.len() - 1;
    auto sr1 = [&](auto& v) { v[i] = v[last]; v.pop_back(); };
    auto sr3 = [&](auto& v) { for (int c = 0; c < 3; ++c) v[3 * i + c] = v[3 * last + c]; v.resize(3 * last); };
    sr1(a.tag); sr1(a.atom_type); sr1(a.origin_index); sr1(a.is_ghost);
    sr3(a.pos); sr3(a.vel); sr3(a.force);
    sr1(a.cutoff_radius); sr1(a.mass); sr1(a.inv_mass); sr3(a.image);
    for (auto& c : b->cols) {
        for (int w = 0; w < c.width; ++w) c.data[i * c.width + w] = c.data[last * c.width + w];
        c.data.resize(last * c.width);
    }
}

// domain.rs:674-729
void pbc_triclinic(soil_ref_box* b) {
    Atom& a = b->atoms;
    const Domain& d = b->domain;
    const double* bvel = d.boundary_vel;
    size_t nlocal_before = a.nlocal;
    size_t removed = 0;
    for (size_t ii = a.nlocal; ii-- > 0;) {
        double pr[3] = {a.pos[3 * ii], a.pos[3 * ii + 1], a.pos[3 * ii + 2]};
        double lam[3];
        d.x2lamda(pr, lam);
        int delta[3] = {0, 0, 0};
        bool gone = false;
        for (int k = 0; k < 3; ++k) {
            if (d.periodic[k]) {
                if (lam[k] < 0.0) { lam[k] += 1.0; delta[k] = -1; }
                else if (lam[k] >= 1.0) { lam[k] -= 1.0; delta[k] = 1; }
            } else if (lam[k] < 0.0 || lam[k] >= 1.0) {
                atom_swap_remove(b, ii);
                removed += 1;
                gone = true;
                break;
            }
        }
        if (gone) continue;
        double x[3];
        d.lamda2x(lam, x);
        a.pos[3 * ii] = x[0]; a.pos[3 * ii + 1] = x[1]; a.pos[3 * ii + 2] = x[2];
        for (int k = 0; k < 3; ++k) a.image[3 * ii + k] += delta[k];
        if (delta[1] != 0) {
            double s = (double)delta[1];
            a.vel[3 * ii] -= s * bvel[0];
            a.vel[3 * ii + 1] -= s * bvel[1];
            a.vel[3 * ii + 2] -= s * bvel[2];
        }
    }
    if (removed > 0) {
        a.nlocal = (uint32_t)(nlocal_before - removed);
        b->state = FullRebuild;
    }
}

// domain.rs:611-663
void pbc(soil_ref_box* b) {
    Atom& a = b->atoms;
    const Domain& d = b->domain;
    if (d.triclinic) { pbc_triclinic(b); return; }
    if (d.periodic[0] && d.periodic[1] && d.periodic[2]) {
        for (size_t i = 0; i < a.nlocal; ++i)
            for (int k = 0; k < 3; ++k) {
                int delta;
                a.pos[3 * i + k] = wrap_periodic(a.pos[3 * i + k], d.boundaries_low[k], d.size[k], &delta);
                a.image[3 * i + k] += delta;
            }
        return;
    }
    size_t nlocal_before = a.nlocal;
    size_t removed = 0;
    for (size_t ii = a.nlocal; ii-- > 0;) {
        for (int k = 0; k < 3; ++k) {
            if (d.periodic[k]) {
                int delta;
                a.pos[3 * ii + k] = wrap_periodic(a.pos[3 * ii + k], d.boundaries_low[k], d.size[k], &delta);
                a.image[3 * ii + k] += delta;
            } else if (a.pos[3 * ii + k] < d.boundaries_low[k] || a.pos[3 * ii + k] >= d.boundaries_high[k]) {
                atom_swap_remove(b, ii);
                removed += 1;
                goto next_atom;
            }
        }
    next_atom:;
    }
    if (removed > 0) {
        a.nlocal = (uint32_t)(nlocal_before - removed);
        b->state = FullRebuild;
    }
}

// atom.rs:525-571 pack_border + unpack_atom as one in-place append (self-send).
void append_ghost(soil_ref_box* b, size_t i, const double cp[3], const double vo[3]) {
    Atom& a = b->atoms;
    a.tag.push_back(a.tag[i]);
    a.origin_index.push_back((int32_t)(double)i);
    a.cutoff_radius.push_back(a.cutoff_radius[i]);
    a.atom_type.push_back(a.atom_type[i]);
    double p0 = a.pos[3 * i] + cp[0], p1 = a.pos[3 * i + 1] + cp[1], p2 = a.pos[3 * i + 2] + cp[2];
    double v0 = a.vel[3 * i] + vo[0], v1 = a.vel[3 * i + 1] + vo[1], v2 = a.vel[3 * i + 2] + vo[2];
    double f0 = a.force[3 * i], f1 = a.force[3 * i + 1], f2 = a.force[3 * i + 2];
    a.pos.push_back(p0); a.pos.push_back(p1); a.pos.push_back(p2);
    a.vel.push_back(v0); a.vel.push_back(v1); a.vel.push_back(v2);
    a.force.push_back(f0); a.force.push_back(f1); a.force.push_back(f2);
    double m = a.mass[i];
    a.mass.push_back(m);
    a.inv_mass.push_back(1.0 / m);
    int32_t i0 = a.image[3 * i], i1 = a.image[3 * i + 1], i2 = a.image[3 * i + 2];
    a.image.push_back(i0); a.image.push_back(i1); a.image.push_back(i2);
    a.is_ghost.push_back(1);
    cols_push_copy(b, i);
    a.nghost += 1;
}

// comm.rs:288-358 (selection) fused with the self-send unpack (comm.rs:794-799)
size_t pack_border_atoms(soil_ref_box* b, int dim, int swap, const double change_pos[3],
                         const double vel_offset[3], std::vector<size_t>& packed, size_t scan_end,
                         double ghost_cutoff, bool self_send) {
    Atom& a = b->atoms;
    const Domain& d = b->domain;
    packed.clear();
    bool use_fixed_cut = ghost_cutoff > 0.0;
    // select first (the scan reads only [0, scan_end), which the appends never touch)
    if (d.triclinic) {
        double lo_face = d.sub_lamda_low[dim], hi_face = d.sub_lamda_high[dim];
        double inflate = dim == 0 ? 1.0 + std::fabs(d.tilt[0]) / d.size[1] : 1.0;
        double cut_l = use_fixed_cut ? (ghost_cutoff / d.size[dim]) * inflate : 0.0;
        for (size_t i = 0; i < scan_end; ++i) {
            double pr[3] = {a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2]};
            double lam[3];
            d.x2lamda(pr, lam);
            double lam_dim = lam[dim];
            bool in_skin;
            if (use_fixed_cut) in_skin = swap == 0 ? lam_dim < lo_face + cut_l : lam_dim >= hi_face - cut_l;
            else {
                double cut = (a.cutoff_radius[i] * 4.0 / d.size[dim]) * inflate;
                in_skin = swap == 0 ? lam_dim < lo_face + cut : lam_dim >= hi_face - cut;
            }
            if (in_skin) packed.push_back(i);
        }
    } else {
        double lo_face = d.sub_domain_low[dim], hi_face = d.sub_domain_high[dim];
        double fixed_thresh = swap == 0 ? lo_face + ghost_cutoff : hi_face - ghost_cutoff;
        for (size_t i = 0; i < scan_end; ++i) {
            double pos_dim = a.pos[3 * i + dim];
            bool in_skin;
            if (use_fixed_cut) in_skin = swap == 0 ? pos_dim < fixed_thresh : pos_dim >= fixed_thresh;
            else {
                double cut = a.cutoff_radius[i] * 4.0;
                in_skin = swap == 0 ? pos_dim < lo_face + cut : pos_dim >= hi_face - cut;
            }
            if (in_skin) packed.push_back(i);
        }
    }
    if (self_send)
        for (size_t i : packed) append_ghost(b, i, change_pos, vel_offset);
    return packed.size();
}

// comm.rs:682-860 (single rank: every periodic swap is a self-send)
void borders(soil_ref_box* b) {
    Atom& a = b->atoms;
    const Domain& d = b->domain;
    CommTopology& topo = b->topo;
    if (a.nghost > 0) {
        atoms_truncate_to_nlocal(a);
        cols_truncate(b, a.nlocal);
        a.nghost = 0;
    }
    a.nlocal = (uint32_t)a.len();
    a.nghost = 0;
    if (topo.maxneed[0] == 0 && topo.maxneed[1] == 0 && topo.maxneed[2] == 0)
        for (int k = 0; k < 3; ++k) topo.maxneed[k] = 1;  // decomp == 1
    size_t swap_idx = 0;
    size_t scan_end = a.nlocal;
    for (int dim = 0; dim < 3; ++dim) {
        for (int need = 0; need < topo.maxneed[dim]; ++need) {
            size_t dim_scan_end = scan_end;
            for (int swap = 0; swap < 2; ++swap) {
                int to_proc = topo.swap_directions[swap][dim];
                int from_proc = topo.swap_directions[(swap + 1) % 2][dim];
                double po = topo.periodic_swap[swap][dim];
                double pbc_shift[3];
                if (dim == 0) { pbc_shift[0] = po * d.size[0]; pbc_shift[1] = 0.0; pbc_shift[2] = 0.0; }
                else if (dim == 1) { pbc_shift[0] = po * d.tilt[0]; pbc_shift[1] = po * d.size[1]; pbc_shift[2] = 0.0; }
                else { pbc_shift[0] = po * d.tilt[1]; pbc_shift[1] = po * d.tilt[2]; pbc_shift[2] = po * d.size[2]; }
                double vel_shift[3] = {0.0, 0.0, 0.0};
                if (dim == 1)
                    for (int k = 0; k < 3; ++k) vel_shift[k] = po * d.boundary_vel[k];
                if (swap_idx == topo.swap_data.size()) topo.swap_data.emplace_back();
                SwapData& e = topo.swap_data[swap_idx];
                bool self_send = to_proc == RANK;
                size_t recv_start = a.len();
                size_t count = pack_border_atoms(b, dim, swap, pbc_shift, vel_shift, e.send_indices,
                                                 dim_scan_end, d.ghost_cutoff, self_send);
                e.recv_start = recv_start;
                e.recv_count = self_send ? count : 0;
                e.to_proc = to_proc;
                e.from_proc = from_proc;
                for (int k = 0; k < 3; ++k) { e.periodic_offset[k] = pbc_shift[k]; e.vel_offset[k] = vel_shift[k]; }
                swap_idx += 1;
            }
            scan_end = (size_t)a.nlocal + a.nghost;
        }
    }
    topo.swap_data.resize(swap_idx);
}

// comm.rs:400-432, 496-573, 582-592 (self-send rounds, applied swap by swap)
void forward_comm_borders(soil_ref_box* b) {
    Atom& a = b->atoms;
    for (const SwapData& s : b->topo.swap_data) {
        if (s.to_proc != RANK) continue;  // non-periodic: nothing received on one rank
        for (size_t k = 0; k < s.recv_count; ++k) {
            size_t idx = s.send_indices[k];
            size_t g = s.recv_start + k;
            // pack reads idx (which may itself be a ghost refreshed earlier), then unpack writes g
            double p0 = a.pos[3 * idx] + s.periodic_offset[0];
            double p1 = a.pos[3 * idx + 1] + s.periodic_offset[1];
            double p2 = a.pos[3 * idx + 2] + s.periodic_offset[2];
            double v0 = a.vel[3 * idx] + s.vel_offset[0];
            double v1 = a.vel[3 * idx + 1] + s.vel_offset[1];
            double v2 = a.vel[3 * idx + 2] + s.vel_offset[2];
            a.pos[3 * g] = p0; a.pos[3 * g + 1] = p1; a.pos[3 * g + 2] = p2;
            a.vel[3 * g] = v0; a.vel[3 * g + 1] = v1; a.vel[3 * g + 2] = v2;
            for (auto& c : b->cols)
                if (c.flags & COL_FORWARD)
                    for (int w = 0; w < c.width; ++w) c.data[g * c.width + w] = c.data[idx * c.width + w];
        }
    }
}

// comm.rs:867-878, 938-1024 (rounds last -> first; within a round lo then hi)
void reverse_send_force(soil_ref_box* b) {
    Atom& a = b->atoms;
    auto& swaps = b->topo.swap_data;
    size_t end = swaps.size();
    auto one = [&](const SwapData& s) {
        if (s.from_proc != RANK) return;  // to_proc == -1 on one rank: nothing accumulated
        // pack ghost forces first (k-th ghost), then accumulate onto send_indices[k]
        std::vector<double> buf;
        buf.reserve(s.recv_count * 3);
        std::vector<std::vector<double>> cbuf(b->cols.size());
        for (size_t i = s.recv_start; i < s.recv_start + s.recv_count; ++i) {
            buf.push_back(a.force[3 * i]); buf.push_back(a.force[3 * i + 1]); buf.push_back(a.force[3 * i + 2]);
            for (size_t ci = 0; ci < b->cols.size(); ++ci)
                if (b->cols[ci].flags & COL_REVERSE)
                    for (int w = 0; w < b->cols[ci].width; ++w)
                        cbuf[ci].push_back(b->cols[ci].data[i * b->cols[ci].width + w]);
        }
        for (size_t k = 0; k < s.send_indices.size(); ++k) {
            size_t o = s.send_indices[k];
            a.force[3 * o] += buf[3 * k];
            a.force[3 * o + 1] += buf[3 * k + 1];
            a.force[3 * o + 2] += buf[3 * k + 2];
            for (size_t ci = 0; ci < b->cols.size(); ++ci) {
                Column& c = b->cols[ci];
                if (c.flags & COL_REVERSE)
                    for (int w = 0; w < c.width; ++w) c.data[o * c.width + w] += cbuf[ci][k * c.width + w];
            }
        }
    };
    while (end > 0) {
        size_t start = end >= 2 ? end - 2 : 0;
        one(swaps[start]);
        if (end > start + 1) one(swaps[start + 1]);
        end = start;
    }
}

// soil_verlet/src/lib.rs:131-162
void initial_integration(soil_ref_box* b) {
    Atom& a = b->atoms;
    double dt_a = a.dt;
    for (size_t i = 0; i < a.nlocal; ++i) {
        double h = 0.5 * dt_a * a.inv_mass[i];
        double v0 = a.vel[3 * i] + h * a.force[3 * i];
        double v1 = a.vel[3 * i + 1] + h * a.force[3 * i + 1];
        double v2 = a.vel[3 * i + 2] + h * a.force[3 * i + 2];
        a.vel[3 * i] = v0; a.vel[3 * i + 1] = v1; a.vel[3 * i + 2] = v2;
        a.pos[3 * i] = a.pos[3 * i] + v0 * dt_a;
        a.pos[3 * i + 1] = a.pos[3 * i + 1] + v1 * dt_a;
        a.pos[3 * i + 2] = a.pos[3 * i + 2] + v2 * dt_a;
    }
}

// soil_verlet/src/lib.rs:175-194
void final_integration(soil_ref_box* b) {
    Atom& a = b->atoms;
    double dt_a = a.dt;
    for (size_t i = 0; i < a.nlocal; ++i) {
        double h = 0.5 * dt_a * a.inv_mass[i];
        a.vel[3 * i] = a.vel[3 * i] + h * a.force[3 * i];
        a.vel[3 * i + 1] = a.vel[3 * i + 1] + h * a.force[3 * i + 1];
        a.vel[3 * i + 2] = a.vel[3 * i + 2] + h * a.force[3 * i + 2];
    }
}

// soil_deform/src/lib.rs:435-604, Lees–Edwards xy branch (:559-603)
// soil_deform/src/lib.rs:435-604: axis styles first (new bounds as functions of
// the step, affine remap of the locals, domain update), then the xy shear.
// Returns false when an axis would collapse (the reference panics).
bool apply_deform(soil_ref_box* b) {
    DeformState& st = b->deform;
    if (!st.initialized || !st.has_any()) return true;
    st.step += 1;
    Atom& a = b->atoms;
    Domain& d = b->domain;
    double dt = a.dt;
    size_t step = st.step;
    size_t nlocal = a.nlocal;
    size_t total_steps = st.stage_steps;
    bool box_changed = false;
    for (int dim = 0; dim < 3; ++dim) {
        const AxisDeform& ax = st.axes[dim];
        if (ax.style < 0) continue;
        double old_lo = d.boundaries_low[dim], old_hi = d.boundaries_high[dim];
        double old_size = old_hi - old_lo;
        double lo_0 = ax.lo_0, hi_0 = ax.hi_0;
        double l0 = hi_0 - lo_0;
        double new_lo, new_hi;
        if (ax.style == 0) {
            double scale = 1.0 + ax.value * dt * (double)step;
            double new_l = l0 * scale;
            double center = (lo_0 + hi_0) * 0.5;
            new_lo = center - new_l * 0.5;
            new_hi = center + new_l * 0.5;
        } else if (ax.style == 1) {
            double delta = ax.value * dt * (double)step;
            double new_l = l0 + delta;
            double center = (lo_0 + hi_0) * 0.5;
            new_lo = center - new_l * 0.5;
            new_hi = center + new_l * 0.5;
        } else {
            double frac = total_steps > 0 ? (double)step / (double)total_steps : 1.0;
            frac = std::fmin(frac, 1.0);
            new_lo = lo_0 + (ax.target_lo - lo_0) * frac;
            new_hi = hi_0 + (ax.target_hi - hi_0) * frac;
        }
        double new_size = new_hi - new_lo;
        if (new_size <= 0.0) return false;
        if (st.remap && old_size > 0.0) {
            double scale = new_size / old_size;
            double new_center = (new_lo + new_hi) * 0.5;
            double old_center = (old_lo + old_hi) * 0.5;
            for (size_t i = 0; i < nlocal; ++i) a.pos[3 * i + dim] = new_center + (a.pos[3 * i + dim] - old_center) * scale;
        }
        d.boundaries_low[dim] = new_lo;
        d.boundaries_high[dim] = new_hi;
        d.size[dim] = new_size;
        d.sub_domain_low[dim] = new_lo;
        d.sub_domain_high[dim] = new_hi;
        d.sub_length[dim] = new_size;
        box_changed = true;
    }
    if (st.has_shear) {
        double rate = st.shear_xy_rate;
        double lx = d.size[0], ly = d.size[1];
        double dv = rate * ly;
        if (!st.shear_initialized) {
            double yc = 0.5 * (d.boundaries_low[1] + d.boundaries_high[1]);
            for (size_t i = 0; i < nlocal; ++i) a.vel[3 * i] += rate * (a.pos[3 * i + 1] - yc);
            st.shear_initialized = true;
        }
        double raw = rate * dt * (double)step * ly;
        double xy = raw - lx * std::round(raw / lx);
        d.tilt[0] = xy;
        d.boundary_vel[0] = dv;
        d.triclinic = true;
        d.bounds_changed = true;
        box_changed = true;
    }
    if (box_changed) {
        d.bounds_changed = true;
        b->neighbor.last_build_pos.clear();
    }
    return true;
}

// Harmonic soft-sphere test consumer (not reference code). For each CSR pair
// (i, j): F = k (ri+rj-d)/d * (xj - xi) for 0 < d < ri+rj, evaluated once in
// f64, then quantized to q = round_half_even(F * 2^32) per component;
// atom i gets -q, atom j gets +q, accumulated in int64 (exact, so the order
// of accumulation cannot matter), and force += acc * 2^-32 (exact).
const double kFixScale = 4294967296.0;  // 2^32

inline void pair_force(const Atom& a, size_t i, size_t j, double k, double F[3]) {
    double dx = a.pos[3 * j] - a.pos[3 * i];
    double dy = a.pos[3 * j + 1] - a.pos[3 * i + 1];
    double dz = a.pos[3 * j + 2] - a.pos[3 * i + 2];
    double r2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
    double rs = a.cutoff_radius[i] + a.cutoff_radius[j];
    if (r2 < rs * rs && r2 > 0.0) {
        double dist = std::sqrt(r2);
        double s = k * (rs - dist) / dist;
        F[0] = s * dx; F[1] = s * dy; F[2] = s * dz;
    } else {
        F[0] = 0.0; F[1] = 0.0; F[2] = 0.0;
    }
}

void harmonic_force(soil_ref_box* b, double k) {
    Atom& a = b->atoms;
    const Neighbor& nb = b->neighbor;
    size_t n_all = a.len();
    size_t nlocal = a.nlocal;
    std::vector<int64_t> acc(3 * n_all, 0);
    for (size_t i = 0; i < nlocal; ++i)
        for (uint32_t e = nb.neighbor_offsets[i]; e < nb.neighbor_offsets[i + 1]; ++e) {
            size_t j = nb.neighbor_indices[e];
            double F[3];
            pair_force(a, i, j, k, F);
            for (int c = 0; c < 3; ++c) {
                int64_t q = (int64_t)std::llrint(F[c] * kFixScale);
                acc[3 * i + c] -= q;
                acc[3 * j + c] += q;
            }
        }
    for (size_t ai = 0; ai < n_all; ++ai)
        for (int c = 0; c < 3; ++c) a.force[3 * ai + c] += (double)acc[3 * ai + c] * (1.0 / kFixScale);
}

// One timestep in the phase order of schedule.rs:22-52 with the plugin
// registrations of neighbor.rs:541-566, comm.rs:177-199, domain.rs:441-449,
// atom.rs:613-619, soil_verlet/src/lib.rs:103-118, soil_deform/src/lib.rs:249-269.
// domain.rs:509-569 shrink_wrap_update: each shrink-wrap axis takes the
// extent of the locals plus a padding (explicit, else the ghost cutoff, else
// 1% of the largest size); a move beyond 1e-12 on either side changes the
// bounds and marks them changed.
bool shrink_wrap(soil_ref_box* b) {
    Domain& d = b->domain;
    const Atom& a = b->atoms;
    const size_t nlocal = a.nlocal;
    if (!(d.shrink[0] || d.shrink[1] || d.shrink[2]) || nlocal == 0) return false;
    double padding;
    if (d.shrink_wrap_padding > 0.0) {
        padding = d.shrink_wrap_padding;
    } else if (d.ghost_cutoff > 0.0) {
        padding = d.ghost_cutoff;
    } else {
        const double m = std::fmax(std::fmax(d.size[0], d.size[1]), d.size[2]);
        padding = m > 0.0 ? m * 0.01 : 1.0;
    }
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
        if (!d.shrink[k]) continue;
        double pmin = DBL_MAX, pmax = -DBL_MAX;
        for (size_t i = 0; i < nlocal; ++i) {
            const double p = a.pos[3 * i + k];
            if (p < pmin) pmin = p;
            if (p > pmax) pmax = p;
        }
        const double nl = pmin - padding, nh = pmax + padding;
        if (std::fabs(nl - d.boundaries_low[k]) > 1e-12 || std::fabs(nh - d.boundaries_high[k]) > 1e-12) {
            d.boundaries_low[k] = nl;
            d.boundaries_high[k] = nh;
            d.sub_domain_low[k] = nl;
            d.sub_domain_high[k] = nh;
            changed = true;
        }
    }
    if (changed) {
        for (int k = 0; k < 3; ++k) {  // update_derived (domain.rs:173-179)
            d.size[k] = d.boundaries_high[k] - d.boundaries_low[k];
            d.sub_length[k] = d.sub_domain_high[k] - d.sub_domain_low[k];
        }
        d.bounds_changed = true;
    }
    return changed;
}

void step_once(soil_ref_box* b) {
    if (!apply_deform(b)) b->collapsed = true;         // PreInitialIntegration
    initial_integration(b);                            // InitialIntegration
    if (b->state == CommunicateOnly) decide_rebuild(b);  // PostInitialIntegration
    if (b->state == FullRebuild) remove_ghost_atoms(b);
    zero_all_forces(b);
    shrink_wrap(b);                                    // PreExchange, before pbc
    if (b->state == FullRebuild) pbc(b);
    sort_atoms_by_bin(b);                              // PreNeighbor
    if (b->state == FullRebuild) borders(b);
    else forward_comm_borders(b);
    if (b->state == FullRebuild) bin_neighbor_list(b); // Neighbor
    if (b->use_consumer) harmonic_force(b, b->harmonic_k);  // Force
    if (b->neighbor.newton) reverse_send_force(b);     // PostForce
    final_integration(b);                              // FinalIntegration
    b->steps += 1;
}

}  // namespace

// ─────────────────────────── C ABI for the tests ────────────────────────────
extern "C" {

soil_ref_box* soil_ref_create(const soil_ref_desc* desc) {
    soil_ref_box* b = new soil_ref_box();
    Domain& d = b->domain;
    for (int k = 0; k < 3; ++k) {
        d.boundaries_low[k] = desc->lo[k];
        d.boundaries_high[k] = desc->hi[k];
        d.sub_domain_low[k] = desc->lo[k];
        d.sub_domain_high[k] = desc->hi[k];
        d.size[k] = desc->hi[k] - desc->lo[k];
        d.sub_length[k] = d.size[k];
        d.periodic[k] = desc->periodic[k];
        d.sub_lamda_low[k] = 0.0;
        d.sub_lamda_high[k] = 1.0;
    }
    Neighbor& nb = b->neighbor;
    nb.skin_fraction = desc->skin_fraction;
    nb.bin_min_size = desc->bin_min_size;
    nb.every = desc->every;
    nb.check = desc->check != 0;
    nb.sort_every = desc->sort_every;
    nb.newton = desc->newton != 0;
    b->atoms.dt = desc->dt;
    b->deform.has_shear = desc->has_shear != 0;
    b->deform.shear_xy_rate = desc->shear_rate;
    return b;
}

void soil_ref_destroy(soil_ref_box* b) { delete b; }

// Append n local atoms (host AoS-of-3 arrays; null vel/force/image = zeros).
int soil_ref_add_atoms(soil_ref_box* b, size_t n, const double* pos, const double* vel, const double* force,
                       const double* mass, const double* radius, const uint32_t* tag, const uint32_t* type,
                       const int32_t* image) {
    Atom& a = b->atoms;
    if (a.nghost != 0) return -1;
    for (size_t i = 0; i < n; ++i) {
        a.tag.push_back(tag ? tag[i] : (uint32_t)(a.len()));
        a.atom_type.push_back(type ? type[i] : 0);
        a.origin_index.push_back(0);
        a.is_ghost.push_back(0);
        for (int k = 0; k < 3; ++k) {
            a.pos.push_back(pos[3 * i + k]);
            a.vel.push_back(vel ? vel[3 * i + k] : 0.0);
            a.force.push_back(force ? force[3 * i + k] : 0.0);
            a.image.push_back(image ? image[3 * i + k] : 0);
        }
        a.mass.push_back(mass[i]);
        a.inv_mass.push_back(1.0 / mass[i]);
        a.cutoff_radius.push_back(radius[i]);
    }
    a.nlocal += (uint32_t)n;
    for (auto& c : b->cols) c.data.resize(a.len() * c.width, 0.0);
    return 0;
}

int soil_ref_register_column(soil_ref_box* b, int width, unsigned flags) {
    Column c;
    c.width = width;
    c.flags = flags;
    c.data.assign(b->atoms.len() * width, 0.0);
    b->cols.push_back(c);
    return (int)b->cols.size() - 1;
}

int soil_ref_set_column(soil_ref_box* b, int col, const double* data, size_t n) {
    Column& c = b->cols[col];
    if (n * c.width > c.data.size()) return -1;
    copy_bytes(c.data.data(), data, n * c.width * sizeof(double));
    return 0;
}

size_t soil_ref_get_column(soil_ref_box* b, int col, double* out, size_t cap) {
    Column& c = b->cols[col];
    size_t n = std::min(cap, c.data.size());
    if (out) copy_bytes(out, c.data.data(), n * sizeof(double));
    return c.data.size();
}

void soil_ref_set_consumer(soil_ref_box* b, int enabled, double k) {
    b->use_consumer = enabled != 0;
    b->harmonic_k = k;
}

// Setup systems: neighbor_setup (PostSetup), comm_setup, setup_deform.
int soil_ref_setup(soil_ref_box* b) {
    comm_setup(b);
    neighbor_setup(b);
    // setup_deform (soil_deform/src/lib.rs:346-420): bounds of the deformed axes at the start
    for (int dim = 0; dim < 3; ++dim)
        if (b->deform.axes[dim].style >= 0) {
            b->deform.axes[dim].lo_0 = b->domain.boundaries_low[dim];
            b->deform.axes[dim].hi_0 = b->domain.boundaries_high[dim];
        }
    b->deform.initialized = true;
    b->deform.step = 0;
    b->deform.shear_initialized = false;
    return 0;
}

// [deform] x/y/z = { style = "erate" (0) | "vel" (1) | "final" (2), ... } (soil_deform/src/lib.rs:98-126)
int soil_ref_set_axis_deform(soil_ref_box* b, int dim, int style, double value, double target_lo, double target_hi) {
    if (dim < 0 || dim > 2 || style < -1 || style > 2) return -1;
    AxisDeform& ax = b->deform.axes[dim];
    ax.style = style;
    ax.value = value;
    ax.target_lo = target_lo;
    ax.target_hi = target_hi;
    return 0;
}
void soil_ref_set_deform_remap(soil_ref_box* b, int remap) { b->deform.remap = remap != 0; }
void soil_ref_set_stage_steps(soil_ref_box* b, uint64_t n) { b->deform.stage_steps = (size_t)n; }
// [domain] boundary per axis: 0 periodic, 1 fixed, 2 shrink-wrap (domain.rs:27-34)
int soil_ref_set_boundary(soil_ref_box* b, int dim, int type) {
    if (dim < 0 || dim > 2 || type < 0 || type > 2) return -1;
    b->domain.periodic[dim] = type == 0 ? 1 : 0;
    b->domain.shrink[dim] = type == 2;
    return 0;
}
void soil_ref_set_shrink_wrap_padding(soil_ref_box* b, double padding) { b->domain.shrink_wrap_padding = padding; }
void soil_ref_get_domain(soil_ref_box* b, double* out9) {
    for (int k = 0; k < 3; ++k) {
        out9[k] = b->domain.boundaries_low[k];
        out9[3 + k] = b->domain.boundaries_high[k];
        out9[6 + k] = b->domain.size[k];
    }
}

int soil_ref_step(soil_ref_box* b, int nsteps) {
    for (int s = 0; s < nsteps && !b->collapsed; ++s) step_once(b);
    return b->collapsed ? -1 : 0;
}

// Phase-by-phase access for unit tests. Ids follow the step order.
int soil_ref_phase(soil_ref_box* b, int phase) {
    switch (phase) {
        case 0: if (!apply_deform(b)) b->collapsed = true; break;
        case 1: initial_integration(b); break;
        case 2: decide_rebuild(b); break;
        case 3: remove_ghost_atoms(b); break;
        case 4: zero_all_forces(b); break;
        case 5: pbc(b); break;
        case 6: sort_atoms_by_bin(b); break;
        case 7: borders(b); break;
        case 8: forward_comm_borders(b); break;
        case 9: bin_neighbor_list(b); break;
        case 10: harmonic_force(b, b->harmonic_k); break;
        case 11: reverse_send_force(b); break;
        case 12: final_integration(b); break;
        case 13: shrink_wrap(b); break;
        default: return -1;
    }
    return 0;
}

int soil_ref_get_state_flag(soil_ref_box* b) { return (int)b->state; }
void soil_ref_set_state_flag(soil_ref_box* b, int s) { b->state = s ? CommunicateOnly : FullRebuild; }

// counts[0..7] = nlocal, nghost, npairs, ncells, rebuilds, steps, nstencil_fwd, nswaps
void soil_ref_counts(soil_ref_box* b, uint64_t* out) {
    out[0] = b->atoms.nlocal;
    out[1] = b->atoms.nghost;
    out[2] = b->neighbor.neighbor_indices.size();
    out[3] = b->neighbor.bin_total_cells;
    out[4] = b->rebuilds;
    out[5] = b->steps;
    out[6] = b->neighbor.bin_stencil_forward.size();
    out[7] = b->topo.swap_data.size();
}

// grid: bin_count[3] (as doubles), bin_size[3], bin_origin[3], ghost_cutoff, max_cutoff, bin_min_size,
//       tilt[3], boundary_vel[3], cached_min_skin, uniform_flag, uniform_cutoff_sq
void soil_ref_grid(soil_ref_box* b, double* out) {
    const Neighbor& nb = b->neighbor;
    const Domain& d = b->domain;
    for (int k = 0; k < 3; ++k) {
        out[k] = nb.bin_count[k];
        out[3 + k] = nb.bin_size[k];
        out[6 + k] = nb.bin_origin[k];
        out[12 + k] = d.tilt[k];
        out[15 + k] = d.boundary_vel[k];
    }
    out[9] = nb.ghost_cutoff;
    out[10] = nb.cached_max_cutoff;
    out[11] = nb.bin_min_size;
    out[18] = nb.cached_min_skin;
    out[19] = nb.has_uniform_cutoff ? 1.0 : 0.0;
    out[20] = nb.cached_uniform_cutoff_sq;
}

size_t soil_ref_stencil(soil_ref_box* b, int forward, int32_t* out, size_t cap) {
    const auto& s = forward ? b->neighbor.bin_stencil_forward : b->neighbor.bin_stencil;
    for (size_t i = 0; i < std::min(cap, s.size()); ++i) out[i] = s[i];
    return s.size();
}

// atom.rs:589-604 compute_ke: sequential sum over the (masked) locals
double soil_ref_compute_ke(soil_ref_box* b, const uint8_t* mask) {
    const Atom& a = b->atoms;
    double ke = 0.0;
    for (size_t i = 0; i < a.nlocal; ++i) {
        if (mask && !mask[i]) continue;
        const double vx = a.vel[3 * i], vy = a.vel[3 * i + 1], vz = a.vel[3 * i + 2];
        ke += a.mass[i] * (vx * vx + vy * vy + vz * vz);
    }
    return 0.5 * ke;
}

// neighbor.rs:1321-1336 classify_boundary_atoms: a local row is "boundary"
// when any of its neighbours is a ghost (index >= nlocal)
void soil_ref_classify_csr(const uint32_t* offsets, const uint32_t* indices, size_t nlocal, uint8_t* out) {
    for (size_t i = 0; i < nlocal; ++i) {
        bool any = false;
        for (uint32_t e = offsets[i]; e < offsets[i + 1]; ++e) any = any || indices[e] >= nlocal;
        out[i] = any ? 1 : 0;
    }
}

void soil_ref_classify_boundary_atoms(soil_ref_box* b, uint8_t* out) {
    if (b->neighbor.neighbor_offsets.size() < (size_t)b->atoms.nlocal + 1) {  // no list built yet
        std::memset(out, 0, b->atoms.nlocal);
        return;
    }
    soil_ref_classify_csr(b->neighbor.neighbor_offsets.data(), b->neighbor.neighbor_indices.data(), b->atoms.nlocal,
                          out);
}

size_t soil_ref_csr(soil_ref_box* b, uint32_t* offsets, uint32_t* indices, size_t cap) {
    const Neighbor& nb = b->neighbor;
    if (offsets) copy_bytes(offsets, nb.neighbor_offsets.data(), nb.neighbor_offsets.size() * 4);
    size_t P = nb.neighbor_indices.size();
    if (indices && cap >= P) copy_bytes(indices, nb.neighbor_indices.data(), P * 4);
    return P;
}

// bins of the last build: bin_start[C+1], sorted_atoms[N+G], atom_sorted_idx[N+G], atom_cell[N+G]
void soil_ref_bins(soil_ref_box* b, uint32_t* bin_start, uint32_t* sorted_atoms, uint32_t* sorted_idx,
                   uint32_t* atom_cell) {
    const Neighbor& nb = b->neighbor;
    if (bin_start) copy_bytes(bin_start, nb.bin_start.data(), nb.bin_start.size() * 4);
    if (sorted_atoms) copy_bytes(sorted_atoms, nb.bin_sorted_atoms.data(), nb.bin_sorted_atoms.size() * 4);
    if (sorted_idx) copy_bytes(sorted_idx, nb.bin_atom_sorted_idx.data(), nb.bin_atom_sorted_idx.size() * 4);
    if (atom_cell) copy_bytes(atom_cell, nb.bin_atom_cell.data(), nb.bin_atom_cell.size() * 4);
}

size_t soil_ref_bins_size(soil_ref_box* b) { return b->neighbor.bin_start.size(); }

// last spatial-sort permutation (perm[k] = old index of new atom k); returns 0 if none ran
size_t soil_ref_sort_perm(soil_ref_box* b, uint64_t* out) {
    if (!b->neighbor.last_sort_ran) return 0;
    const auto& p = b->neighbor.last_sort_perm;
    if (out) copy_bytes(out, p.data(), p.size() * 8);
    return p.size();
}

// all atoms (locals then ghosts)
void soil_ref_atoms(soil_ref_box* b, double* pos, double* vel, double* force, double* radius, double* mass,
                    double* inv_mass, uint32_t* tag, int32_t* origin, int32_t* image) {
    const Atom& a = b->atoms;
    size_t n = a.len();
    if (pos) copy_bytes(pos, a.pos.data(), 3 * n * 8);
    if (vel) copy_bytes(vel, a.vel.data(), 3 * n * 8);
    if (force) copy_bytes(force, a.force.data(), 3 * n * 8);
    if (radius) copy_bytes(radius, a.cutoff_radius.data(), n * 8);
    if (mass) copy_bytes(mass, a.mass.data(), n * 8);
    if (inv_mass) copy_bytes(inv_mass, a.inv_mass.data(), n * 8);
    if (tag) copy_bytes(tag, a.tag.data(), n * 4);
    if (origin) copy_bytes(origin, a.origin_index.data(), n * 4);
    if (image) copy_bytes(image, a.image.data(), 3 * n * 4);
}

// swap table: per swap [recv_start, recv_count, nsend, to_proc, from_proc, off xyz, voff xyz]
size_t soil_ref_swaps(soil_ref_box* b, double* out) {
    const auto& sw = b->topo.swap_data;
    for (size_t s = 0; s < sw.size(); ++s) {
        double* o = out + 11 * s;
        o[0] = (double)sw[s].recv_start; o[1] = (double)sw[s].recv_count; o[2] = (double)sw[s].send_indices.size();
        o[3] = sw[s].to_proc; o[4] = sw[s].from_proc;
        for (int k = 0; k < 3; ++k) { o[5 + k] = sw[s].periodic_offset[k]; o[8 + k] = sw[s].vel_offset[k]; }
    }
    return sw.size();
}

size_t soil_ref_swap_send(soil_ref_box* b, int s, uint64_t* out) {
    const auto& v = b->topo.swap_data[s].send_indices;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return v.size();
}

void soil_ref_set_tilt(soil_ref_box* b, double xy, double bvel_x) {
    b->domain.tilt[0] = xy;
    b->domain.boundary_vel[0] = bvel_x;
    b->domain.triclinic = true;
    b->domain.bounds_changed = true;
}

void soil_ref_set_ghost_cutoff(soil_ref_box* b, double gc) { b->domain.ghost_cutoff = gc; }

// Exposed helpers for the known-answer tests.
void soil_ref_x2lamda(soil_ref_box* b, const double* r, double* out) { b->domain.x2lamda(r, out); }
void soil_ref_lamda2x(soil_ref_box* b, const double* l, double* out) { b->domain.lamda2x(l, out); }
void soil_ref_set_vel(soil_ref_box* b, size_t i, const double* v) {
    for (int k = 0; k < 3; ++k) b->atoms.vel[3 * i + k] = v[k];
}
void soil_ref_set_force(soil_ref_box* b, size_t i, const double* f) {
    for (int k = 0; k < 3; ++k) b->atoms.force[3 * i + k] = f[k];
}
void soil_ref_set_pos(soil_ref_box* b, size_t i, const double* p) {
    for (int k = 0; k < 3; ++k) b->atoms.pos[3 * i + k] = p[k];
}

// Timed CPU baseline: nsteps full steps, wall seconds returned.
double soil_ref_time_steps(soil_ref_box* b, int nsteps);

}  // extern "C"

#include <chrono>
extern "C" double soil_ref_time_steps(soil_ref_box* b, int nsteps) {
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) step_once(b);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
