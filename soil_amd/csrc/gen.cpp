// Synthetic inputs for the five benchmark workloads (host only).
//
// SplitMix64 -> U[0,1) as (u >> 11) * 2^-53, normals by Box–Muller (one normal
// per pair of uniforms). Draw order per box: radii, blob centres, positions,
// velocities. The same seed gives the same box on any host with the same libm.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "soil_hip.h"
#include "internal.hpp"

namespace soil_hip {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        double u1 = u01(), u2 = u01();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
    }
};

static const double kPi = 3.141592653589793;

// Fills radii/positions/velocities for one box and returns the edge lengths.
static void gen_box_atoms(SplitMix64& rng, size_t n, double rmin, double rmax, double phi, const double aspect[3],
                          int nblobs, double blob_frac, double sigma, double vel_sigma, double L[3], double* pos,
                          double* vel, double* mass, double* radius) {
    std::vector<double> r(n);
    double vol = 0.0;
    for (size_t i = 0; i < n; ++i) {
        r[i] = rmax > rmin ? rmin + (rmax - rmin) * rng.u01() : rmin;
        vol += (4.0 / 3.0) * kPi * r[i] * r[i] * r[i];
    }
    double V = vol / phi;
    double a = std::cbrt(V / (aspect[0] * aspect[1] * aspect[2]));
    for (int k = 0; k < 3; ++k) L[k] = a * aspect[k];
    std::vector<double> centres(3 * (size_t)(nblobs > 0 ? nblobs : 0));
    for (int bl = 0; bl < nblobs; ++bl)
        for (int k = 0; k < 3; ++k) centres[3 * bl + k] = L[k] * rng.u01();
    for (size_t i = 0; i < n; ++i) {
        bool in_blob = nblobs > 0 && rng.u01() < blob_frac;
        if (in_blob) {
            int bl = (int)(rng.u01() * nblobs);
            if (bl >= nblobs) bl = nblobs - 1;
            for (int k = 0; k < 3; ++k) {
                double x = centres[3 * bl + k] + sigma * rng.normal();
                x = x - L[k] * std::floor(x / L[k]);
                if (x >= L[k]) x = 0.0;
                pos[3 * i + k] = x;
            }
        } else {
            for (int k = 0; k < 3; ++k) pos[3 * i + k] = L[k] * rng.u01();
        }
    }
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) vel[3 * i + k] = vel_sigma * rng.normal();
    for (size_t i = 0; i < n; ++i) {
        mass[i] = 1.0;
        radius[i] = r[i];
    }
}

static void default_desc(soil_hip_box_desc* d) {
    std::memset(d, 0, sizeof(*d));
    for (int k = 0; k < 3; ++k) d->periodic[k] = 1;
    d->newton = 1;
    d->skin_fraction = 1.1;
    d->bin_min_size = 1.0;
    d->every = 0;
    d->check = 1;
    d->sort_every = 1000;
    d->dt = 1e-3;
}

}  // namespace soil_hip

using namespace soil_hip;

extern "C" int soil_hip_gen_custom(uint64_t seed, size_t n, double rmin, double rmax, double phi, const double* aspect,
                                   int nblobs, double blob_frac, double sigma, double vel_sigma, double shear_rate,
                                   soil_hip_box_desc* desc, double* pos, double* vel, double* mass, double* radius) {
    if (!desc || !pos || !vel || !mass || !radius || n == 0 || phi <= 0.0 || rmin <= 0.0 || rmax < rmin)
        return set_error(SOIL_HIP_EINVAL, "gen_custom: bad arguments");
    SplitMix64 rng(seed);
    double L[3];
    double asp[3] = {1.0, 1.0, 1.0};
    if (aspect)
        for (int k = 0; k < 3; ++k) asp[k] = aspect[k];
    gen_box_atoms(rng, n, rmin, rmax, phi, asp, nblobs, blob_frac, sigma, vel_sigma, L, pos, vel, mass, radius);
    default_desc(desc);
    for (int k = 0; k < 3; ++k) {
        desc->lo[k] = 0.0;
        desc->hi[k] = L[k];
    }
    desc->has_shear = shear_rate != 0.0 ? 1 : 0;
    desc->shear_rate = shear_rate;
    return SOIL_HIP_OK;
}

// Per-config draws of box parameters (counts, packing, shear) come from a
// separate stream seeded by (seed, box_id) so that a box can be regenerated alone.
static void config_params(int config, int box_id, uint64_t seed, size_t* n, double* rmin, double* rmax, double* phi,
                          double aspect[3], int* nblobs, double* blob_frac, double* sigma, double* shear) {
    SplitMix64 p(seed ^ 0xA5A5A5A5ull ^ ((uint64_t)(box_id + 1) * 0x9E3779B97F4A7C15ull));
    aspect[0] = aspect[1] = aspect[2] = 1.0;
    *nblobs = 0;
    *blob_frac = 0.0;
    *sigma = 0.0;
    *shear = 0.0;
    switch (config) {
        case 1:
            *n = 4000; *rmin = *rmax = 0.5; *phi = 0.30;
            break;
        case 2:
            *n = 1000000; *rmin = 0.4; *rmax = 0.6; *phi = 0.50;
            break;
        case 3:
            *n = 4000000; *rmin = 0.45; *rmax = 0.55; *phi = 0.40;
            aspect[0] = 3.0;
            *nblobs = 64; *blob_frac = 0.5; *sigma = 4.0; *shear = 0.1;
            break;
        case 4: {
            double lnN = std::log(1e3) + (std::log(64e3) - std::log(1e3)) * p.u01();
            *n = (size_t)std::llround(std::exp(lnN));
            *phi = 0.20 + 0.35 * p.u01();
            *rmin = 0.4; *rmax = 0.6;
            *shear = 0.2 * p.u01();
            if (*shear == 0.0) *shear = 1e-6;
            break;
        }
        case 5: {
            double lnN = std::log(2e3) + (std::log(200e3) - std::log(2e3)) * p.u01();
            *n = (size_t)std::llround(std::exp(lnN));
            *phi = 0.20 + 0.35 * p.u01();
            *rmin = 0.4; *rmax = 0.6;
            bool le = p.u01() < 0.5;
            double rate = 0.01 + 0.19 * p.u01();
            *shear = le ? rate : 0.0;
            break;
        }
        default:
            *n = 0;
    }
}

extern "C" int soil_hip_gen_box(int config, int box_id, uint64_t seed, soil_hip_box_desc* desc, size_t* n,
                                double* pos, double* vel, double* mass, double* radius) {
    if (config < 1 || config > 5 || !desc || !n) return set_error(SOIL_HIP_EINVAL, "gen_box: config must be 1..5");
    size_t nn;
    double rmin, rmax, phi, aspect[3], blob_frac, sigma, shear;
    int nblobs;
    config_params(config, box_id, seed, &nn, &rmin, &rmax, &phi, aspect, &nblobs, &blob_frac, &sigma, &shear);
    *n = nn;
    if (!pos) {
        // size query: the description still needs the box lengths, so generate radii only
        SplitMix64 rng(seed + (uint64_t)config + (uint64_t)box_id);
        double vol = 0.0;
        for (size_t i = 0; i < nn; ++i) {
            double r = rmax > rmin ? rmin + (rmax - rmin) * rng.u01() : rmin;
            vol += (4.0 / 3.0) * kPi * r * r * r;
        }
        double a = std::cbrt((vol / phi) / (aspect[0] * aspect[1] * aspect[2]));
        default_desc(desc);
        for (int k = 0; k < 3; ++k) desc->hi[k] = a * aspect[k];
        desc->has_shear = shear != 0.0;
        desc->shear_rate = shear;
        return SOIL_HIP_OK;
    }
    return soil_hip_gen_custom(seed + (uint64_t)config + (uint64_t)box_id, nn, rmin, rmax, phi, aspect, nblobs,
                               blob_frac, sigma, 0.1, shear, desc, pos, vel, mass, radius);
}

// Served-stream job metadata: steps ~ U{500..5000}; arrival times of a Poisson
// process of rate `rate` (jobs per second), all from one seeded stream.
extern "C" int soil_hip_gen_arrivals(uint64_t seed, int njobs, double rate, double* arrival_s, int32_t* steps) {
    if (njobs < 0 || rate <= 0.0) return set_error(SOIL_HIP_EINVAL, "gen_arrivals: bad arguments");
    SplitMix64 rng(seed ^ 0x5EEDull);
    double t = 0.0;
    for (int j = 0; j < njobs; ++j) {
        t += -std::log(1.0 - rng.u01()) / rate;
        if (arrival_s) arrival_s[j] = t;
        int s = 500 + (int)(rng.u01() * 4501.0);
        if (s > 5000) s = 5000;
        if (steps) steps[j] = s;
    }
    return SOIL_HIP_OK;
}
