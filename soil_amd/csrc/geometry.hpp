// Host model of a box's geometry (see geometry.cpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "internal.hpp"
#include "soil_hip.h"

namespace soil_hip {

struct HostModel {
    // Domain (domain.rs:105-142), one rank: sub-domain = box
    double lo[3], hi[3], size[3], sub_lo[3], sub_hi[3], sub_len[3];
    bool periodic[3];
    bool shrink[3];  // shrink-wrap axes (non-periodic), domain.rs:27-34
    double shrink_padding;
    double ghost_cutoff;
    double tilt[3], bvel[3];
    bool triclinic, bounds_changed;
    // Neighbor (neighbor.rs:87-174)
    double skin_fraction, bin_min_size;
    uint64_t every;
    bool check, newton;
    uint64_t sort_every, sort_counter;
    double cached_max_cutoff = 0.0, cached_min_skin = 0.0, ucut_sq = 0.0;
    bool uniform = false;
    int nbin[3];
    double bin_size[3], bin_origin[3], pbc_box[3];
    uint64_t ncells = 0;
    std::vector<int> stencil_fwd, stencil_full;
    bool has_self = false;
    // Atom::dt and DeformState (soil_deform/src/lib.rs:168-211): per-axis style
    // (-1 none, 0 erate, 1 vel, 2 final), parameters and bounds at setup
    double dt;
    bool has_shear, shear_initialized, deform_initialized;
    double shear_rate;
    uint64_t deform_step;
    int axis_style[3];
    double axis_value[3], axis_target_lo[3], axis_target_hi[3], axis_lo0[3], axis_hi0[3];
    bool remap;
    uint64_t stage_steps;
    // this step's affine remap of the locals, per axis (consumed by fill())
    bool ev_remap[3];
    double remap_old_c[3], remap_new_c[3], remap_scale[3];
    // events of the current step (consumed by fill())
    bool ev_init_shear, ev_clear_last, ev_bounds_changed, ev_periodic_sort;

    void init(const soil_hip_box_desc& d);
    void setup(const double* radius, size_t n);
    void compute_grid();
    bool has_deform() const;
    bool has_shrink() const { return shrink[0] || shrink[1] || shrink[2]; }
    // shrink_wrap from the locals' extents (min[3], max[3]); true when the bounds moved
    bool shrink_wrap_update(const double* pmin, const double* pmax, size_t nlocal);
    bool apply_deform();  // false when an axis collapses
    void set_tilt(double xy, double bvel_x);
    void sort_prepare();
    int fill(Geom& g) const;
    uint64_t max_cells() const;
    size_t max_stencil() const;
    size_t ghost_estimate(size_t nlocal) const;
};

}  // namespace soil_hip
