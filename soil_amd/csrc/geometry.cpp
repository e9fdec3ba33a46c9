// Host model of a box's geometry and per-step schedule.
//
// The bin grid, stencil, Lees–Edwards tilt and periodic-sort counter are
// functions of the box description, the atoms' radii and the step count only,
// never of positions. The host therefore advances this model one step ahead of
// the device and hands each step its Geom record; everything that depends on
// positions (rebuild trigger, bins, ghosts, CSR) is decided on the device.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "geometry.hpp"

namespace soil_hip {

// neighbor.rs:311-397 (orthogonal) — grid, ghost layers, min-distance stencil
static void grid_orthogonal(HostModel& m, double max_cutoff, double min_bin) {
    int ni[3];
    double bs[3];
    for (int k = 0; k < 3; ++k) {
        ni[k] = (int)std::max(std::floor(m.sub_len[k] / min_bin), 1.0);
        bs[k] = m.sub_len[k] / (double)ni[k];
    }
    double ghost_cut = m.ghost_cutoff > max_cutoff ? m.ghost_cutoff : max_cutoff;
    int s[3];
    for (int k = 0; k < 3; ++k) s[k] = (int)std::ceil(ghost_cut / bs[k]) + (int)std::ceil(max_cutoff / bs[k]);
    for (int k = 0; k < 3; ++k) {
        m.nbin[k] = ni[k] + 2 * s[k];
        m.bin_size[k] = bs[k];
        m.bin_origin[k] = m.sub_lo[k] - bs[k] * (double)s[k];
    }
    const int ny = m.nbin[1], nz = m.nbin[2];
    const double cutoff_sq = max_cutoff * max_cutoff;
    m.stencil_fwd.clear();
    m.stencil_full.clear();
    m.has_self = false;
    for (int dx = -s[0]; dx <= s[0]; ++dx)
        for (int dy = -s[1]; dy <= s[1]; ++dy)
            for (int dz = -s[2]; dz <= s[2]; ++dz) {
                const double mx = (double)std::max(std::abs(dx) - 1, 0) * bs[0];
                const double my = (double)std::max(std::abs(dy) - 1, 0) * bs[1];
                const double mz = (double)std::max(std::abs(dz) - 1, 0) * bs[2];
                if (mx * mx + my * my + mz * mz < cutoff_sq) {
                    const int off = dx * ny * nz + dy * nz + dz;
                    m.stencil_full.push_back(off);
                    if (off > 0) m.stencil_fwd.push_back(off);
                    else if (off == 0) m.has_self = true;
                }
            }
    for (int k = 0; k < 3; ++k) m.pbc_box[k] = m.hi[k] - m.lo[k];
}

// neighbor.rs:411-495 (triclinic): lamda-space grid, tilt-dependent stencil
static void grid_triclinic(HostModel& m, double max_cutoff, double min_bin) {
    const double* l = m.size;
    const double xy = m.tilt[0];
    const double sub_l[3] = {1.0 - 0.0, 1.0 - 0.0, 1.0 - 0.0};
    int ni[3];
    double bl[3], rb[3];
    for (int k = 0; k < 3; ++k) {
        ni[k] = (int)std::max(std::floor(sub_l[k] * l[k] / min_bin), 1.0);
        bl[k] = sub_l[k] / (double)ni[k];
        rb[k] = bl[k] * l[k];
    }
    const double ghost_cut = m.ghost_cutoff > max_cutoff ? m.ghost_cutoff : max_cutoff;
    int sr[3];
    for (int k = 0; k < 3; ++k) sr[k] = (int)std::ceil(max_cutoff / rb[k]);
    const int skew = (int)std::ceil(std::fabs(xy) / rb[0]);
    const int sx = (int)std::ceil(ghost_cut / rb[0]) + sr[0] + skew;
    const int sy = (int)std::ceil(ghost_cut / rb[1]) + sr[1];
    const int sz = (int)std::ceil(ghost_cut / rb[2]) + sr[2];
    m.nbin[0] = ni[0] + 2 * sx;
    m.nbin[1] = ni[1] + 2 * sy;
    m.nbin[2] = ni[2] + 2 * sz;
    for (int k = 0; k < 3; ++k) m.bin_size[k] = bl[k];
    m.bin_origin[0] = 0.0 - bl[0] * (double)sx;
    m.bin_origin[1] = 0.0 - bl[1] * (double)sy;
    m.bin_origin[2] = 0.0 - bl[2] * (double)sz;
    const int ny = m.nbin[1], nz = m.nbin[2];
    const double cutoff_sq = max_cutoff * max_cutoff;
    m.stencil_fwd.clear();
    m.stencil_full.clear();
    m.has_self = false;
    for (int dx = -sx; dx <= sx; ++dx)
        for (int dy = -sy; dy <= sy; ++dy)
            for (int dz = -sz; dz <= sz; ++dz) {
                const double m1 = (double)std::max(std::abs(dy) - 1, 0) * bl[1];
                const double m2 = (double)std::max(std::abs(dz) - 1, 0) * bl[2];
                const double dyz_sq = (l[1] * m1) * (l[1] * m1) + (l[2] * m2) * (l[2] * m2);
                if (dyz_sq >= cutoff_sq) continue;
                const double budget = std::sqrt(cutoff_sq - dyz_sq);
                const double a_lo = l[0] * ((double)dx - 1.0) * bl[0];
                const double a_hi = l[0] * ((double)dx + 1.0) * bl[0];
                const double b1 = xy * ((double)dy - 1.0) * bl[1];
                const double b2 = xy * ((double)dy + 1.0) * bl[1];
                const double b_lo = b1 <= b2 ? b1 : b2;
                const double b_hi = b1 <= b2 ? b2 : b1;
                const double lo = a_lo + b_lo, hi = a_hi + b_hi;
                if (hi >= -budget && lo <= budget) {
                    const int off = dx * ny * nz + dy * nz + dz;
                    m.stencil_full.push_back(off);
                    if (off > 0) m.stencil_fwd.push_back(off);
                    else if (off == 0) m.has_self = true;
                }
            }
    for (int k = 0; k < 3; ++k) m.pbc_box[k] = l[k];
}

void HostModel::compute_grid() {
    const double max_cutoff = cached_max_cutoff;
    const double min_bin = std::max(bin_min_size, max_cutoff);  // one rank: required_bin = max_cutoff
    if (triclinic) grid_triclinic(*this, max_cutoff, min_bin);
    else grid_orthogonal(*this, max_cutoff, min_bin);
    ncells = (uint64_t)nbin[0] * (uint64_t)nbin[1] * (uint64_t)nbin[2];
}

void HostModel::init(const soil_hip_box_desc& d) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = d.lo[k];
        hi[k] = d.hi[k];
        size[k] = d.hi[k] - d.lo[k];
        sub_lo[k] = d.lo[k];
        sub_hi[k] = d.hi[k];
        sub_len[k] = size[k];
        periodic[k] = d.periodic[k] != 0;
        shrink[k] = false;
        tilt[k] = 0.0;
        bvel[k] = 0.0;
    }
    triclinic = false;
    bounds_changed = false;
    shrink_padding = 0.0;
    ghost_cutoff = 0.0;
    skin_fraction = d.skin_fraction;
    bin_min_size = d.bin_min_size;
    every = d.every;
    check = d.check != 0;
    sort_every = d.sort_every;
    sort_counter = 0;
    newton = d.newton != 0;
    dt = d.dt;
    has_shear = d.has_shear != 0;
    shear_rate = d.shear_rate;
    shear_initialized = false;
    deform_step = 0;
    deform_initialized = false;
    for (int k = 0; k < 3; ++k) {
        axis_style[k] = -1;
        axis_value[k] = axis_target_lo[k] = axis_target_hi[k] = axis_lo0[k] = axis_hi0[k] = 0.0;
        ev_remap[k] = false;
        remap_old_c[k] = remap_new_c[k] = remap_scale[k] = 0.0;
    }
    remap = true;
    stage_steps = 1;
    ev_init_shear = ev_clear_last = ev_bounds_changed = ev_periodic_sort = false;
}

// neighbor_setup (neighbor.rs:604-665) from the uploaded radii, plus the
// per-build constants of save_build_positions (neighbor.rs:669-690), which on
// a fixed atom set depend only on the radii.
void HostModel::setup(const double* radius, size_t n) {
    double max_skin = 0.0, min_skin = DBL_MAX;
    for (size_t i = 0; i < n; ++i) {
        max_skin = std::fmax(max_skin, radius[i]);
        min_skin = std::fmin(min_skin, radius[i]);
    }
    const double max_cutoff = max_skin > 0.0 ? 2.0 * max_skin * skin_fraction : bin_min_size;
    const double disp_buf = min_skin < DBL_MAX * 0.5 ? (skin_fraction - 1.0) * min_skin : 0.0;
    const double gc = std::fmax(max_cutoff + 2.0 * disp_buf, ghost_cutoff);
    ghost_cutoff = gc;
    cached_max_cutoff = max_cutoff;
    if (bin_min_size < max_cutoff) bin_min_size = max_cutoff;
    compute_grid();
    // save_build_positions constants (locals only; a fixed set)
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (size_t i = 0; i < n; ++i) {
        mn = std::fmin(mn, radius[i]);
        mx = std::fmax(mx, radius[i]);
    }
    cached_min_skin = mn;
    uniform = std::fabs(mx - mn) < 1e-15;
    if (uniform) {
        const double c = 2.0 * mn * skin_fraction;
        ucut_sq = c * c;
    } else {
        ucut_sq = 0.0;
    }
    // setup_deform (soil_deform/src/lib.rs:346-420): the deformed axes' bounds at the start
    for (int k = 0; k < 3; ++k)
        if (axis_style[k] >= 0) {
            axis_lo0[k] = lo[k];
            axis_hi0[k] = hi[k];
        }
    deform_initialized = true;
    deform_step = 0;
    shear_initialized = false;
}

// shrink_wrap_update (domain.rs:509-569), with the extents of the locals
// reduced on the device (min and max are exact in any order)
bool HostModel::shrink_wrap_update(const double* pmin, const double* pmax, size_t nlocal) {
    if (!has_shrink() || nlocal == 0) return false;
    double padding;
    if (shrink_padding > 0.0) {
        padding = shrink_padding;
    } else if (ghost_cutoff > 0.0) {
        padding = ghost_cutoff;
    } else {
        const double m = std::fmax(std::fmax(size[0], size[1]), size[2]);
        padding = m > 0.0 ? m * 0.01 : 1.0;
    }
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
        if (!shrink[k]) continue;
        const double nl = pmin[k] - padding, nh = pmax[k] + padding;
        if (std::fabs(nl - lo[k]) > 1e-12 || std::fabs(nh - hi[k]) > 1e-12) {
            lo[k] = nl;
            hi[k] = nh;
            sub_lo[k] = nl;
            sub_hi[k] = nh;
            changed = true;
        }
    }
    if (changed) {
        for (int k = 0; k < 3; ++k) {  // update_derived (domain.rs:173-179)
            size[k] = hi[k] - lo[k];
            sub_len[k] = sub_hi[k] - sub_lo[k];
        }
        bounds_changed = true;
    }
    return changed;
}

bool HostModel::has_deform() const {
    return has_shear || axis_style[0] >= 0 || axis_style[1] >= 0 || axis_style[2] >= 0;
}

// apply_deform (soil_deform/src/lib.rs:435-604): axis styles (new bounds as a
// function of the step, the locals' affine remap recorded for the device, the
// domain updated), then the xy Lees–Edwards tilt.
bool HostModel::apply_deform() {
    ev_init_shear = false;
    ev_clear_last = false;
    for (int k = 0; k < 3; ++k) ev_remap[k] = false;
    if (!deform_initialized || !has_deform()) return true;
    deform_step += 1;
    const uint64_t step = deform_step;
    bool box_changed = false;
    for (int dim = 0; dim < 3; ++dim) {
        if (axis_style[dim] < 0) continue;
        const double old_lo = lo[dim], old_hi = hi[dim];
        const double old_size = old_hi - old_lo;
        const double lo_0 = axis_lo0[dim], hi_0 = axis_hi0[dim];
        const double l0 = hi_0 - lo_0;
        double new_lo, new_hi;
        if (axis_style[dim] == 0) {
            const double scale = 1.0 + axis_value[dim] * dt * (double)step;
            const double new_l = l0 * scale;
            const double center = (lo_0 + hi_0) * 0.5;
            new_lo = center - new_l * 0.5;
            new_hi = center + new_l * 0.5;
        } else if (axis_style[dim] == 1) {
            const double delta = axis_value[dim] * dt * (double)step;
            const double new_l = l0 + delta;
            const double center = (lo_0 + hi_0) * 0.5;
            new_lo = center - new_l * 0.5;
            new_hi = center + new_l * 0.5;
        } else {
            double frac = stage_steps > 0 ? (double)step / (double)stage_steps : 1.0;
            frac = std::fmin(frac, 1.0);
            new_lo = lo_0 + (axis_target_lo[dim] - lo_0) * frac;
            new_hi = hi_0 + (axis_target_hi[dim] - hi_0) * frac;
        }
        const double new_size = new_hi - new_lo;
        if (new_size <= 0.0) return false;
        if (remap && old_size > 0.0) {
            ev_remap[dim] = true;
            remap_scale[dim] = new_size / old_size;
            remap_new_c[dim] = (new_lo + new_hi) * 0.5;
            remap_old_c[dim] = (old_lo + old_hi) * 0.5;
        }
        lo[dim] = new_lo;
        hi[dim] = new_hi;
        size[dim] = new_size;
        sub_lo[dim] = new_lo;
        sub_hi[dim] = new_hi;
        sub_len[dim] = new_size;
        box_changed = true;
    }
    if (has_shear) {
        const double lx = size[0], ly = size[1];
        const double dv = shear_rate * ly;
        if (!shear_initialized) {
            ev_init_shear = true;
            shear_initialized = true;
        }
        const double raw = shear_rate * dt * (double)step * ly;
        const double xy = raw - lx * std::round(raw / lx);
        tilt[0] = xy;
        bvel[0] = dv;
        triclinic = true;
        box_changed = true;
    }
    if (box_changed) {
        bounds_changed = true;
        ev_clear_last = true;
    }
    return true;
}

void HostModel::set_tilt(double xy, double bvel_x) {
    tilt[0] = xy;
    bvel[0] = bvel_x;
    triclinic = true;
    bounds_changed = true;
}

// sort_atoms_by_bin bookkeeping (neighbor.rs:836-848)
void HostModel::sort_prepare() {
    ev_bounds_changed = false;
    if (bounds_changed) {
        if (cached_max_cutoff > 0.0) compute_grid();
        bounds_changed = false;
        ev_bounds_changed = true;
    }
    sort_counter += 1;
    ev_periodic_sort = sort_every > 0 && sort_counter >= sort_every;
    if (ev_periodic_sort) sort_counter = 0;
}

int HostModel::fill(Geom& g) const {
    std::memset(&g, 0, sizeof(g));
    for (int k = 0; k < 3; ++k) {
        g.lo[k] = lo[k];
        g.hi[k] = hi[k];
        g.size[k] = size[k];
        g.tilt[k] = tilt[k];
        g.bvel[k] = bvel[k];
        g.bin_size[k] = bin_size[k];
        g.inv_bin[k] = 1.0 / bin_size[k];
        g.bin_origin[k] = bin_origin[k];
        g.pbc_box[k] = pbc_box[k];
        g.inv_pbc_box[k] = 1.0 / pbc_box[k];
        g.half_pbc_box[k] = pbc_box[k] * 0.5;
        g.nbin[k] = nbin[k];
        g.periodic[k] = periodic[k] ? 1 : 0;
    }
    g.ncells = (uint32_t)ncells;
    g.triclinic = triclinic ? 1 : 0;
    g.pbc_all = (periodic[0] && periodic[1] && periodic[2]) ? 1 : 0;
    // border thresholds (comm.rs:305-346): lamda faces of the unit cube or real faces
    for (int d = 0; d < 3; ++d) {
        if (triclinic) {
            const double inflate = d == 0 ? 1.0 + std::fabs(tilt[0]) / size[1] : 1.0;
            const double cut_l = (ghost_cutoff / size[d]) * inflate;
            g.thr_lo[d] = 0.0 + cut_l;
            g.thr_hi[d] = 1.0 - cut_l;
        } else {
            g.thr_lo[d] = sub_lo[d] + ghost_cutoff;
            g.thr_hi[d] = sub_hi[d] - ghost_cutoff;
        }
        g.disjoint[d] = g.thr_lo[d] <= g.thr_hi[d] ? 1 : 0;
    }
    // swap shifts (comm.rs:264-281 signs, 749-761 shifts)
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            const double po = periodic[d] ? (s == 0 ? 1.0 : -1.0) : 0.0;
            double* sh = g.swap_shift[2 * d + s];
            double* vs = g.swap_vshift[2 * d + s];
            if (d == 0) { sh[0] = po * size[0]; sh[1] = 0.0; sh[2] = 0.0; }
            else if (d == 1) { sh[0] = po * tilt[0]; sh[1] = po * size[1]; sh[2] = 0.0; }
            else { sh[0] = po * tilt[1]; sh[1] = po * tilt[2]; sh[2] = po * size[2]; }
            for (int k = 0; k < 3; ++k) vs[k] = d == 1 ? po * bvel[k] : 0.0;
        }
    g.dt = dt;
    g.shear_rate = shear_rate;
    g.yc = 0.5 * (lo[1] + hi[1]);
    for (int k = 0; k < 3; ++k) {
        g.remap[k] = ev_remap[k] ? 1 : 0;
        g.remap_any |= g.remap[k];
        g.remap_old_c[k] = remap_old_c[k];
        g.remap_new_c[k] = remap_new_c[k];
        g.remap_scale[k] = remap_scale[k];
    }
    g.skin_fraction = skin_fraction;
    g.sf_sq = skin_fraction * skin_fraction;
    const double pas = (skin_fraction - 1.0) * cached_min_skin;
    g.disp_thr_sq = pas * pas;
    g.ucut_sq = ucut_sq;
    g.uniform = uniform ? 1 : 0;
    g.ghost_cut = ghost_cutoff;
    g.every = every;
    g.check = check ? 1 : 0;
    g.newton = newton ? 1 : 0;
    g.has_self = has_self ? 1 : 0;
    if (newton) {
        if (stencil_fwd.size() > SOIL_MAX_STENCIL) return -1;
        g.nstencil = (int32_t)stencil_fwd.size();
        for (size_t k = 0; k < stencil_fwd.size(); ++k) g.stencil[k] = stencil_fwd[k];
    } else {
        int n = 0;
        for (int off : stencil_full) {
            if (off == 0) continue;
            if (n >= SOIL_MAX_STENCIL) return -1;
            g.stencil[n++] = off;
        }
        g.nstencil = n;
    }
    {
        // ascending-offset order of the stencil entries plus the self cell (entry nstencil)
        std::vector<std::pair<int, int>> o;
        for (int k = 0; k < g.nstencil; ++k) o.push_back({g.stencil[k], k});
        o.push_back({0, g.nstencil});
        std::stable_sort(o.begin(), o.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            return a.first < b.first;
        });
        for (size_t k = 0; k < o.size(); ++k) g.sorder[k] = (int16_t)o[k].second;
    }
    g.nruns = 0;
    for (int k = 0; k < g.nstencil; ++k) {
        if (g.nruns > 0 && g.stencil[k] == g.run_off[g.nruns - 1] + g.run_len[g.nruns - 1]) {
            g.run_len[g.nruns - 1] += 1;
        } else {
            g.run_off[g.nruns] = g.stencil[k];
            g.run_len[g.nruns] = 1;
            g.nruns += 1;
        }
    }
    g.init_shear = ev_init_shear ? 1 : 0;
    g.clear_last = ev_clear_last ? 1 : 0;
    g.bounds_changed = ev_bounds_changed ? 1 : 0;
    g.periodic_sort = ev_periodic_sort ? 1 : 0;
    return 0;
}

// Largest cell count over every tilt the box can reach (box flip bounds |xy| to Lx/2).
uint64_t HostModel::max_cells() const {
    if (!has_shear && !triclinic) return ncells;
    HostModel m = *this;
    m.triclinic = true;
    uint64_t best = ncells;
    const double xs[3] = {0.0, 0.5 * size[0], -0.5 * size[0]};
    for (double xy : xs) {
        m.tilt[0] = xy;
        m.compute_grid();
        best = std::max(best, m.ncells);
    }
    return best;
}

size_t HostModel::max_stencil() const {
    HostModel m = *this;
    size_t best = newton ? stencil_fwd.size() : stencil_full.size();
    if (!has_shear && !triclinic) return best;
    m.triclinic = true;
    for (int k = 0; k <= 20; ++k) {
        m.tilt[0] = -0.5 * size[0] + size[0] * (double)k / 20.0;
        m.compute_grid();
        best = std::max(best, newton ? m.stencil_fwd.size() : m.stencil_full.size());
    }
    return best;
}

// Ghost-count estimate for capacity: shell volume of the ghost cut around the
// box (x inflated by the largest tilt the box can reach), times the density.
size_t HostModel::ghost_estimate(size_t nlocal) const {
    double f = 1.0;
    for (int d = 0; d < 3; ++d) {
        if (!periodic[d]) continue;
        double gc = ghost_cutoff;
        if (d == 0 && has_shear) gc *= 1.0 + 0.5 * size[0] / size[1];
        f *= 1.0 + 2.0 * gc / size[d];
    }
    return (size_t)((f - 1.0) * (double)nlocal);
}

}  // namespace soil_hip
