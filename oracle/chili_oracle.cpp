// chili_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of cellular_raza's chili mechanics step, written from the
// reference source text.  It is the parity checker for the HIP backend and the
// timed CPU baseline; nothing in the product path links or calls it.  Only the
// test suite, the smoke check and the benchmark's cpu_baseline leg load it.
//
// What it restates (paths relative to the cellular_raza workspace):
//   grid / voxel index / stencil     building-blocks/src/domains/cartesian_cuboid_n.rs:19-59,183-256,320-337,469-491,586-662
//   subdomain split                  cartesian_cuboid_n.rs:493-584 (get_decomp_res 19-59)
//   initial placement                concepts-derive/src/domain.rs:292-421, core/src/backend/chili/datastructures.rs:278-299
//   pair loop + weights              core/src/backend/chili/update_mechanics.rs:63-109,153-238,285-382
//   boundary + re-binning            update_mechanics.rs:387-513, cartesian_cuboid_n.rs:594-631
//   integrators + AB ring            chili/solvers.rs:23-347, chili/aux_storage.rs:41-130
//   potentials                       building-blocks/src/cell_building_blocks/interaction.rs:84-182,184-299,301-434,497-509
//   mechanics                        mechanics.rs:14-226,228-387,389-571
//   rods                             bacterial_rods.rs:146-337,489-555
//   cell-sorting interaction         benchmarks/src/cell_sorting.rs:31-94, examples/cell_sorting/src/main.rs:32-103,
//                                    examples/cell_sorting/src/main-langevin.rs:44-92
//   main loop order                  core-proc-macro/src/run_sim.rs:220-409
//
// Arithmetic is restated operation by operation (no contraction: this file is
// built with -ffp-contract=off), because rustc does not fuse multiply-adds.
//
// Parity pins (tests/): closed forms and special cases of every potential, the
// reference doctests and unit tests (nearest point, grid, decomposition,
// stencil, Euler / AB bounds), free-drift and harmonic-spring closed forms,
// Langevin MSD statistics, brute-force voxel assignment, and Philox KAT vectors.
// The pair weighting and the order in which a full ring buffer is iterated are
// pinned by reading the reference code and by structural tests derived from it
// (tests/test_oracle_kats.py: same-voxel / adjacent-voxel / cross-subdomain
// weights, the AB recursion under both iteration orders); no reference test or
// fixture holds their values, so beyond those checks they are parity unpinned.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "crmech.h"


namespace orc {

// ----------------------------------------------------------------------------
// Grid: CartesianCuboid (cartesian_cuboid_n.rs:183-256)
// ----------------------------------------------------------------------------
struct Grid {
    int D = 0;
    double min[3] = {0, 0, 0}, max[3] = {0, 0, 0}, dx[3] = {1, 1, 1};
    uint64_t n[3] = {1, 1, 1};
    uint64_t V = 1;
};

// num-traits ToPrimitive::to_usize for f64: truncation, valid for -1 < x < 2^64, NaN -> None.
static bool to_usize(double x, uint64_t& out) {
    if (std::isnan(x)) return false;
    if (!(x > -1.0)) return false;
    if (!(x < 18446744073709551616.0)) return false;
    double t = std::trunc(x);
    out = (t <= 0.0) ? 0 : (uint64_t)t;
    return true;
}

static crm_status make_grid(const crm_domain& d, int D, Grid& g, std::string& err) {
    g.D = D;
    for (int i = 0; i < D; i++) {
        if (d.min[i] >= d.max[i]) {  // check_min_max (cartesian_cuboid_n.rs:149-162)
            err = "BoundaryError: Min must be smaller than Max for domain boundaries";
            return CRM_BOUNDARY_ERROR;
        }
    }
    g.V = 1;
    for (int i = 0; i < 3; i++) {
        g.min[i] = i < D ? d.min[i] : 0.0;
        g.max[i] = i < D ? d.max[i] : 1.0;
        g.n[i] = 1;
        g.dx[i] = 1.0;
    }
    if (d.n_voxels[0] == 0) {
        // from_boundaries_and_interaction_range
        for (int i = 0; i < D; i++) {
            double n = std::floor((d.max[i] - d.min[i]) / d.interaction_range);
            uint64_t m;
            if (!to_usize(n, m)) {
                err = "BoundaryError: Cannot convert float to usize (conversion error during domain setup)";
                return CRM_BOUNDARY_ERROR;
            }
            g.n[i] = std::max<uint64_t>(m, 1);
            g.dx[i] = (d.max[i] - d.min[i]) / n;
        }
    } else {
        // from_boundaries_and_n_voxels
        for (int i = 0; i < D; i++) {
            if (d.n_voxels[i] == 0) {
                err = "BoundaryError: n_voxels must be positive";
                return CRM_BOUNDARY_ERROR;
            }
            g.n[i] = d.n_voxels[i];
            g.dx[i] = (d.max[i] - d.min[i]) / (double)d.n_voxels[i];
        }
    }
    for (int i = 0; i < D; i++) g.V *= g.n[i];
    return CRM_OK;
}

static inline uint64_t plain_of(const Grid& g, const uint64_t* idx) {
    uint64_t p = 0;
    for (int i = 0; i < g.D; i++) p = p * g.n[i] + idx[i];
    return p;
}

static inline void idx_of(const Grid& g, uint64_t plain, uint64_t* idx) {
    for (int i = g.D - 1; i >= 0; i--) {
        idx[i] = plain % g.n[i];
        plain /= g.n[i];
    }
}

// CartesianSubDomain::get_index_of (cartesian_cuboid_n.rs:469-491).
// Returns 0 ok, 1 conversion failure (BoundaryError).
static int index_of(const Grid& g, const double* p, uint64_t* idx) {
    for (int i = 0; i < g.D; i++) {
        double nv = (p[i] - g.min[i]) / g.dx[i];
        if (!to_usize(nv, idx[i])) return 1;
    }
    return 0;
}

static bool in_grid(const Grid& g, const uint64_t* idx) {
    for (int i = 0; i < g.D; i++)
        if (idx[i] >= g.n[i]) return false;
    return true;
}

// get_neighbor_voxel_indices (cartesian_cuboid_n.rs:641-662): lexicographic order,
// last axis fastest, self excluded.  Lexicographic == ascending plain index.
static void neighbors_of(const Grid& g, uint64_t plain, std::vector<uint64_t>& out) {
    out.clear();
    uint64_t idx[3] = {0, 0, 0};
    idx_of(g, plain, idx);
    uint64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    for (int i = 0; i < g.D; i++) {
        lo[i] = idx[i] == 0 ? 0 : idx[i] - 1;
        hi[i] = std::min<uint64_t>(idx[i] + 2, g.n[i]);
    }
    uint64_t q[3];
    for (q[0] = lo[0]; q[0] < hi[0]; q[0]++)
        for (q[1] = lo[1]; q[1] < hi[1]; q[1]++)
            for (q[2] = lo[2]; q[2] < hi[2]; q[2]++) {
                uint64_t pl = plain_of(g, q);
                if (pl != plain) out.push_back(pl);
            }
}

// get_decomp_res (cartesian_cuboid_n.rs:19-59)
static bool decomp_res(uint64_t n_voxel, uint64_t n_regions, uint64_t& rn, uint64_t& rm, uint64_t& ravg) {
    int64_t average_len = (int64_t)std::ceil((double)n_voxel / (double)n_regions);
    int64_t n = (int64_t)n_regions, m = 0;
    for (uint64_t it = 0; it < n_regions; it++) {
        int64_t residue = (int64_t)n_voxel - average_len * n - (average_len - 1) * m;
        if (residue == 0) {
            rn = (uint64_t)n;
            rm = (uint64_t)m;
            ravg = (uint64_t)average_len;
            return true;
        } else if (residue > 0) {
            if (n == (int64_t)n_regions) {
                average_len += 1;
                n = (int64_t)n_regions;
                m = 0;
            }
        } else {
            n -= 1;
            m += 1;
        }
    }
    return false;
}

// ----------------------------------------------------------------------------
// Small vector helpers restating nalgebra's sequential sums.
// ----------------------------------------------------------------------------
static inline double norm_sq(const double* z, int D) {
    double r = 0.0;
    for (int k = 0; k < D; k++) r = r + z[k] * z[k];
    return r;
}
static inline double dot(const double* a, const double* b, int D) {
    double r = 0.0;
    for (int k = 0; k < D; k++) r = r + a[k] * b[k];
    return r;
}

// ----------------------------------------------------------------------------
// Point potentials.  Return 0 ok, 1 CalcError.
// ----------------------------------------------------------------------------
// calculate_morse_interaction (interaction.rs:151-182)
static void morse(int D, const double* own, const double* ext, double r_own, double r_ext, double cutoff,
                  double strength, double stiff, double* f_own, double* f_ext) {
    double z[3];
    for (int k = 0; k < D; k++) z[k] = own[k] - ext[k];
    double dist = std::sqrt(norm_sq(z, D));
    if (dist > cutoff || dist == 0.0) {
        for (int k = 0; k < D; k++) f_own[k] = f_ext[k] = 0.0;
        return;
    }
    double r = r_own + r_ext;
    double e = std::exp((-stiff) * (dist - r));
    double force = ((((-2.0) * strength) * stiff) * e) * (1.0 - e);
    for (int k = 0; k < D; k++) {
        double dir = z[k] / dist;
        f_own[k] = dir * force;
        f_ext[k] = (-dir) * force;
    }
}

// MiePotential::calculate_force_between (interaction.rs:345-376)
static int mie(int D, const crm_potential& P, const double* own, const double* ext, double* f_own,
               double* f_ext) {
    double z[3];
    for (int k = 0; k < D; k++) z[k] = own[k] - ext[k];
    double r = std::sqrt(norm_sq(z, D));
    if (r == 0.0) return 1;
    if (r > P.cutoff) {
        for (int k = 0; k < D; k++) f_own[k] = f_ext[k] = 0.0;
        return 0;
    }
    double x = P.radius + P.radius;
    double sigma = std::pow(P.em / P.en, 1.0 / (P.en - P.em)) * x;
    double mie_constant = P.en / (P.en - P.em) * std::pow(P.en / P.em, P.em / (P.en - P.em));
    double potential_part =
        P.en / sigma * std::pow(sigma / r, P.en + 1.0) - P.em / sigma * std::pow(sigma / r, P.em + 1.0);
    double force = std::fmin(mie_constant * P.strength * potential_part, P.bound);
    for (int k = 0; k < D; k++) {
        double dir = z[k] / r;
        f_own[k] = dir * force;
        f_ext[k] = (-dir) * force;
    }
    return 0;
}

// BoundLennardJones::calculate_force_between (interaction.rs:122-138)
static void bound_lj(int D, const crm_potential& P, const double* own, const double* ext, double* f_own,
                     double* f_ext) {
    double z[3];
    for (int k = 0; k < D; k++) z[k] = own[k] - ext[k];
    double r = std::sqrt(norm_sq(z, D));
    double val = 4.0 * P.epsilon / r *
                 (12.0 * std::pow(P.sigma / r, 11.0) - 6.0 * std::pow(P.sigma / r, 5.0));
    double mx = P.bound / r;
    double q = (P.cutoff >= r) ? 1.0 : 0.0;
    double m = std::fmin(mx, val);
    for (int k = 0; k < D; k++) {
        double dir = z[k] / r;
        f_own[k] = ((-dir) * q) * m;
        f_ext[k] = (dir * q) * m;
    }
}

// CellSpecificInteraction (benchmarks/src/cell_sorting.rs:31-94 with sign = +1,
// examples/cell_sorting/src/main.rs:32-103 with sign = -1).  Both agents share the
// interaction parameters; the ext info is (cell_radius, species).
static void cell_sorting(int D, const crm_potential& P, const double* own, const double* ext, int sp_own,
                         int sp_ext, double* f_own, double* f_ext) {
    const double mrd = 0.3162277660168379;
    double z[3];
    for (int k = 0; k < D; k++) z[k] = own[k] - ext[k];
    double nz = std::sqrt(norm_sq(z, D));
    double r, dir[3];
    if (!(nz < P.radius * mrd)) {
        r = nz;
        for (int k = 0; k < D; k++) dir[k] = z[k] / nz;
    } else {
        bool same = true;
        for (int k = 0; k < D; k++) same = same && (own[k] == ext[k]);
        if (same) {
            for (int k = 0; k < D; k++) f_own[k] = f_ext[k] = 0.0;
            return;
        }
        for (int k = 0; k < D; k++) dir[k] = z[k] / nz;
        r = P.radius * mrd;
    }
    double ext_radius = P.radius;
    double sigma = r / (P.radius + ext_radius);
    double bound = 4.0 + 1.0 / sigma;
    double sgn_arg = P.relative_interaction_range * (P.radius + ext_radius) - r;
    double signum = std::isnan(sgn_arg) ? sgn_arg : std::copysign(1.0, sgn_arg);
    double spatial_cutoff = (1.0 + signum) * 0.5;
    double inv = 1.0 / sigma;
    double strength =
        P.strength * std::fmax(std::fmin(std::pow(inv, 2.0) - std::pow(inv, 4.0), bound), -bound);
    double smax = std::fmax(strength, 0.0), smin = std::fmin(strength, 0.0);
    for (int k = 0; k < D; k++) {
        double att = (dir[k] * smax) * spatial_cutoff;
        double rep = (dir[k] * smin) * spatial_cutoff;
        double o, e;
        if (sp_own == sp_ext) {
            o = rep + att;
            e = (-rep) - att;
        } else {
            o = rep;
            e = -rep;
        }
        if (P.sign < 0) {
            f_own[k] = e;
            f_ext[k] = o;
        } else {
            f_own[k] = o;
            f_ext[k] = e;
        }
    }
}

// CellSpecificInteraction of examples/cell_sorting/src/main-langevin.rs:44-92.
// Both agents share the parameters (cell_radius, repulsion = P.strength, cutoff);
// the attraction of the (own, ext) species pair is species_matrix[own][ext]
// (the example's match on (own, ext), 66-72).  Returns (-force, force) (89).
static void sorting_langevin(int D, const crm_potential& P, const double* own, const double* ext, int sp_own,
                             int sp_ext, double* f_own, double* f_ext) {
    const double own_radius = P.radius, ext_radius = P.radius;
    const double radius_combined = own_radius + ext_radius;
    double z[3], dir[3];
    for (int k = 0; k < D; k++) z[k] = own[k] - ext[k];
    const double r = std::sqrt(norm_sq(z, D));
    const double attraction = P.species_matrix[sp_own * CRM_MAX_SPECIES + sp_ext];
    for (int k = 0; k < D; k++) dir[k] = r == 0.0 ? 0.0 : z[k] / r;  // normalize = z / |z|
    double force[3];
    if (r <= radius_combined) {
        const double s = P.strength * (radius_combined - r);
        for (int k = 0; k < D; k++) force[k] = s * dir[k];
    } else if (r <= P.cutoff) {
        const double sigma = r / radius_combined;
        const double s = -attraction * (r - radius_combined) * (P.cutoff - r) / P.cutoff *
                         std::exp(-1.0 * std::pow(sigma - 1.0, 2.0));
        for (int k = 0; k < D; k++) force[k] = s * dir[k];
    } else {
        for (int k = 0; k < D; k++) force[k] = 0.0 * dir[k];
    }
    for (int k = 0; k < D; k++) {
        f_own[k] = -force[k];
        f_ext[k] = force[k];
    }
}

static int point_force(int D, const crm_potential& P, const double* own, const double* ext, int sp_own,
                       int sp_ext, double* f_own, double* f_ext) {
    switch (P.kind) {
        case CRM_POTENTIAL_MORSE:
            morse(D, own, ext, P.radius, P.radius, P.cutoff, P.strength, P.potential_stiffness, f_own, f_ext);
            return 0;
        case CRM_POTENTIAL_MORSE_SPECIES: {
            double s = P.strength * P.species_matrix[sp_own * CRM_MAX_SPECIES + sp_ext];
            morse(D, own, ext, P.radius, P.radius, P.cutoff, s, P.potential_stiffness, f_own, f_ext);
            return 0;
        }
        case CRM_POTENTIAL_MIE:
            return mie(D, P, own, ext, f_own, f_ext);
        case CRM_POTENTIAL_BOUND_LENNARD_JONES:
            bound_lj(D, P, own, ext, f_own, f_ext);
            return 0;
        case CRM_POTENTIAL_CELL_SORTING:
            cell_sorting(D, P, own, ext, sp_own, sp_ext, f_own, f_ext);
            return 0;
        case CRM_POTENTIAL_NONE:  // NoInteraction (interaction.rs:12-35)
            for (int k = 0; k < D; k++) f_own[k] = f_ext[k] = 0.0;
            return 0;
        case CRM_POTENTIAL_CELL_SORTING_LANGEVIN:
            sorting_langevin(D, P, own, ext, sp_own, sp_ext, f_own, f_ext);
            return 0;
    }
    return 1;
}

// nearest_point_from_point_to_line (interaction.rs:497-509)
static void nearest_point(int D, const double* p, const double* a, const double* b, double& dist, double* x,
                          double& t) {
    double ab[3], ap[3];
    for (int k = 0; k < D; k++) {
        ab[k] = b[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    double tt = dot(ab, ap, D) / norm_sq(ab, D);
    // f64::clamp(0, 1): NaN stays NaN
    if (tt < 0.0) tt = 0.0;
    if (tt > 1.0) tt = 1.0;
    t = tt;
    double d[3];
    for (int k = 0; k < D; k++) {
        x[k] = a[k] * (1.0 - tt) + b[k] * tt;
        d[k] = p[k] - x[k];
    }
    dist = std::sqrt(norm_sq(d, D));
}

// RodInteraction::calculate_force_between (bacterial_rods.rs:292-336), point
// potential = the configured one (Morse for the rods workload).
static int rod_force(int D, int nv, const crm_potential& P, const double* own, const double* ext, int sp_own,
                     int sp_ext, double* f_own, double* f_ext) {
    for (int i = 0; i < nv * D; i++) f_own[i] = f_ext[i] = 0.0;
    for (int i = 0; i < nv; i++) {
        const double* p1 = own + i * D;
        for (int j = 0; j + 1 < nv; j++) {
            double dist, x[3], t;
            nearest_point(D, p1, ext + j * D, ext + (j + 1) * D, dist, x, t);
            double fo[3], fe[3];
            if (point_force(D, P, p1, x, sp_own, sp_ext, fo, fe)) return 1;
            int j1 = (j + 1) % nv;
            for (int k = 0; k < D; k++) {
                f_own[i * D + k] = f_own[i * D + k] + fo[k];
                f_ext[j * D + k] = f_ext[j * D + k] + fe[k] * (1.0 - t);
                f_ext[j1 * D + k] = f_ext[j1 * D + k] + fe[k] * t;
            }
        }
    }
    return 0;
}

// nalgebra Matrix::angle
static double angle(const double* a, const double* b, int D) {
    double prod = dot(a, b, D);
    double n1 = std::sqrt(norm_sq(a, D)), n2 = std::sqrt(norm_sq(b, D));
    if (n1 == 0.0 || n2 == 0.0) return 0.0;
    double cang = prod / (n1 * n2);
    if (cang < -1.0) cang = -1.0;
    if (cang > 1.0) cang = 1.0;
    return std::acos(cang);
}

// RodMechanics::calculate_increment (bacterial_rods.rs:157-213): returns (vel, F_total).
static void rod_increment(int D, int nv, const crm_mechanics& M, const double* pos, const double* vel,
                          const double* force, double* dx, double* dv) {
    const double one_half = 0.5, two = 2.0;
    double total[CRM_MAX_ROD_VERTICES * 3];
    for (int i = 0; i < nv * D; i++) total[i] = force[i];
    double dist[CRM_MAX_ROD_VERTICES * 3];
    for (int i = 0; i + 1 < nv; i++)
        for (int k = 0; k < D; k++) dist[i * D + k] = pos[i * D + k] - pos[(i + 1) * D + k];
    for (int i = 0; i + 1 < nv; i++) {
        const double* d = dist + i * D;
        double nd = std::sqrt(norm_sq(d, D));
        if (nd != 0.0) {
            for (int k = 0; k < D; k++) {
                double dir = d[k] / nd;
                double fi = ((-dir) * M.spring_tension) * (nd - M.spring_length);
                total[i * D + k] = total[i * D + k] + fi * one_half;
                total[(i + 1) * D + k] = total[(i + 1) * D + k] + (-fi) * one_half;
            }
        }
    }
    for (int i = 0; i + 2 < nv; i++) {
        const double* c1 = dist + i * D;
        const double* c2 = dist + (i + 1) * D;
        double ang = angle(c1, c2, D);
        double fd[3];
        for (int k = 0; k < D; k++) fd[k] = c1[k] - c2[k];
        double nfd = std::sqrt(norm_sq(fd, D));
        double tn = std::tan(one_half * ang);
        for (int k = 0; k < D; k++) {
            double fdir = (nfd != 0.0) ? fd[k] / nfd : fd[k];
            double f = ((fdir * two) * M.rigidity) * tn;
            total[i * D + k] = total[i * D + k] + (-f) * one_half;
            total[(i + 1) * D + k] = total[(i + 1) * D + k] + f;
            total[(i + 2) * D + k] = total[(i + 2) * D + k] + (-f) * one_half;
        }
    }
    for (int i = 0; i < nv * D; i++) {
        dv[i] = total[i] - vel[i] * M.damping;
        dx[i] = vel[i];
    }
}

// ----------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11) and Box-Muller.  The same stream is generated on the GPU.
// ----------------------------------------------------------------------------
static void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// n standard normals for (seed, agent, step): block b uses counter
// {agent_lo, agent_hi, step_lo, (step_hi << 16) | b}; each block yields two normals.
static void gaussians(uint64_t seed, uint64_t agent, uint64_t step, int n, double* out) {
    const double two_pi = 6.283185307179586;
    const double inv53 = 1.0 / 9007199254740992.0;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; 2 * b < n; b++) {
        uint32_t ctr[4] = {(uint32_t)agent, (uint32_t)(agent >> 32), (uint32_t)step,
                           (uint32_t)(((step >> 32) << 16) | (uint64_t)b)};
        uint32_t r[4];
        philox4x32_10(ctr, key, r);
        uint64_t a = ((uint64_t)r[1] << 32) | r[0];
        uint64_t c = ((uint64_t)r[3] << 32) | r[2];
        double u1 = (double)((a >> 11) + 1) * inv53;  // (0, 1]
        double u2 = (double)(c >> 11) * inv53;        // [0, 1)
        double rad = std::sqrt(-2.0 * std::log(u1));
        double th = two_pi * u2;
        out[2 * b] = rad * std::cos(th);
        if (2 * b + 1 < n) out[2 * b + 1] = rad * std::sin(th);
    }
}

// ----------------------------------------------------------------------------
// Engine.  KM = doubles stored per position (3 for points, 48 for rods with up
// to 16 vertices in 3-D), so point agents stay compact.
// ----------------------------------------------------------------------------
constexpr int KMAX = CRM_MAX_ROD_VERTICES * 3;

template <int KM>
struct Cell {
    uint64_t id;  // agent index
    int32_t species;
    double pos[KM];
    double vel[KM];
    // AuxStorageMechanics (aux_storage.rs:41-130)
    double force[KM];
    double ring_dx[2][KM];
    double ring_dv[2][KM];
    int ring_size;  // entries stored, 0..cap
    int ring_head;  // index of the oldest entry
    // rods: RodMechanics::spring_length, per agent (bacterial_rods.rs:82-107)
    double spring_length;
    // Cycle of the rods example (bacterial_rods/src/main.rs:48-111)
    double growth_rate;
    uint32_t sensed;       // NeighborSensing accumulator of the current step
    bool divide;           // CycleEvent::Division pending
    uint64_t ident[3];     // CellIdentifier: kind (0 Initial(a), 1 Division(a = voxel, b = counter)), a, b
    uint64_t parent[3];    // kind 3: no parent
};

template <int KM>
struct PosInfo {
    double pos[KM], vel[KM];
    int32_t species;
    uint64_t cell_index;
    uint64_t sender, receiver;
};
template <int KM>
struct ForceInfo {
    double force[KM];
    uint64_t cell_index;
    uint64_t sender;
};
template <int KM>
struct SendCell {
    uint64_t sender;
    Cell<KM> cell;
};

struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int n, count = 0;
    uint64_t gen = 0;
    explicit Barrier(int n_) : n(n_) {}
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        uint64_t g = gen;
        if (++count == n) {
            count = 0;
            gen++;
            cv.notify_all();
        } else {
            cv.wait(lk, [&] { return gen != g; });
        }
    }
};

struct StepError {
    std::atomic<int> code{0};
    std::mutex m;
    std::string msg;
    void set(int c, const std::string& s) {
        int expected = 0;
        if (code.compare_exchange_strong(expected, c)) {
            std::lock_guard<std::mutex> lk(m);
            msg = s;
        }
    }
};

struct SimBase {
    crm_config cfg;
    int D = 2, nv = 1, K = 2;
    int cap = 2;  // ring capacity = solver order (0, 1, 2)
    uint32_t n_sub = 1;
    uint64_t n_agents = 0;
    uint64_t step = 0;
    std::string err;
    bool have_agents = false;
    const double* noise = nullptr;  // current step's noise slice
    virtual ~SimBase() {}
    virtual int init(const crm_instance* instances, uint32_t n_instances) = 0;
    virtual int set_agents(const uint64_t* counts, const double* pos, const double* vel,
                           const int32_t* species) = 0;
    virtual int run_phase(bool forces_only) = 0;
    virtual void forces_out(double* force) = 0;
    virtual void zero_forces() = 0;
    virtual void state_out(double* pos, double* vel) = 0;
    virtual uint32_t history_out(double* dx, double* dv) = 0;
    virtual void voxels_out(uint32_t* voxel_of_agent, uint32_t* start_out, uint64_t* order_out) = 0;
    virtual uint64_t n_voxels_total() = 0;
    virtual void neighbor_counts(double radius, uint32_t* counts) = 0;
    virtual int set_domain_force(uint32_t ii, const crm_domain_force& f) = 0;
    virtual int set_rod_cycle(const crm_rod_cycle& y) = 0;
    virtual int set_spring_lengths(const double* sl) = 0;
    virtual void cycle_out(double* spring_length, double* growth_rate, uint64_t* ident, uint64_t* parent) = 0;
    virtual int remove_agents(const uint8_t* flags) = 0;
    virtual int insert_agents(uint64_t n, const uint64_t* mother, const double* mpos, const double* mvel,
                              const double* dpos, const double* dvel, const int32_t* dsp) = 0;
    bool cycle_on = false;
    crm_rod_cycle cyc{};
};

template <int KM>
struct Sim : SimBase {
    struct Instance {
        Grid grid;
        crm_potential pot;
        crm_mechanics mech;
        crm_domain_force dforce{};                      // kind NONE unless set
        std::vector<std::vector<Cell<KM>>> voxels;      // indexed by plain index
        std::vector<uint64_t> id_counter;               // Voxel::id_counter per plain index
        std::vector<uint32_t> owner;                    // subdomain of each voxel
        std::vector<std::vector<uint64_t>> sub_voxels;  // ascending plain indices per subdomain
    };
    struct Inbox {
        std::mutex m;
        std::vector<PosInfo<KM>> pos;
        std::vector<ForceInfo<KM>> force;
        std::vector<SendCell<KM>> cells;
    };
    std::vector<Instance> inst;
    std::vector<std::vector<Inbox>> inbox;  // [instance][subdomain]

    void ring_push(Cell<KM>& c, const double* dx, const double* dv) {
        if (cap == 0) return;
        int slot;
        if (c.ring_size < cap) {
            slot = (c.ring_head + c.ring_size) % cap;
            c.ring_size++;
        } else {
            slot = c.ring_head;
            c.ring_head = (c.ring_head + 1) % cap;
        }
        std::memcpy(c.ring_dx[slot], dx, sizeof(double) * K);
        std::memcpy(c.ring_dv[slot], dv, sizeof(double) * K);
    }

    // i-th element yielded by the ring iterator
    int ring_slot(const Cell<KM>& c, int i) const {
        if (cfg.history_order == CRM_HISTORY_NEWEST_FIRST) return (c.ring_head + c.ring_size - 1 - i) % cap;
        return (c.ring_head + i) % cap;
    }

    int pair_eval(const Instance& I, const Cell<KM>& own, const double* own_pos, const double* ext_pos,
                  int32_t ext_species, double* f_own, double* f_ext) const {
        if (nv == 1) return point_force(D, I.pot, own_pos, ext_pos, own.species, ext_species, f_own, f_ext);
        return rod_force(D, nv, I.pot, own_pos, ext_pos, own.species, ext_species, f_own, f_ext);
    }

    // Voxel::calculate_force_between_cells_external (update_mechanics.rs:63-109).
    int force_external(const Instance& I, std::vector<Cell<KM>>& cells, const double* ext_pos,
                       int32_t ext_species, double* force, bool& has_force) {
        has_force = false;
        double f1[KM], f2[KM];
        for (auto& c : cells) {
            if (pair_eval(I, c, c.pos, ext_pos, ext_species, f1, f2)) return 1;
            for (int k = 0; k < K; k++) c.force[k] = c.force[k] + f1[k] * 0.5;
            if (has_force) {
                for (int k = 0; k < K; k++) force[k] = f2[k] * 0.5 + force[k];
            } else {
                for (int k = 0; k < K; k++) force[k] = f2[k] * 0.5;
                has_force = true;
            }
        }
        return 0;
    }

    // update_mechanics_interaction_step_1 (update_mechanics.rs:153-238)
    int step1(size_t ii, uint32_t sub) {
        Instance& I = inst[ii];
        std::vector<uint64_t> nb;
        double f1[KM], f2[KM], fo[KM], acc[KM];
        for (uint64_t v : I.sub_voxels[sub]) {
            auto& cells = I.voxels[v];
            size_t nc = cells.size();
            for (size_t n = 0; n < nc; n++) {
                Cell<KM>& c1 = cells[n];
                for (size_t m = n + 1; m < nc; m++) {
                    Cell<KM>& c2 = cells[m];
                    if (pair_eval(I, c1, c1.pos, c2.pos, c2.species, f1, f2)) return 1;
                    for (int k = 0; k < K; k++) {
                        c1.force[k] = c1.force[k] + f1[k] * 0.5;
                        c2.force[k] = c2.force[k] + f2[k] * 0.5;
                    }
                    if (pair_eval(I, c2, c2.pos, c1.pos, c1.species, f2, f1)) return 1;
                    for (int k = 0; k < K; k++) {
                        c1.force[k] = c1.force[k] + f1[k] * 0.5;
                        c2.force[k] = c2.force[k] + f2[k] * 0.5;
                    }
                }
                neighbors_of(I.grid, v, nb);
                bool any = false;
                for (uint64_t w : nb) {
                    if (I.owner[w] == sub) {
                        bool has;
                        if (force_external(I, I.voxels[w], c1.pos, c1.species, fo, has)) return 1;
                        if (has) {
                            if (any) {
                                for (int k = 0; k < K; k++) acc[k] = fo[k] * 1.0 + acc[k];
                            } else {
                                for (int k = 0; k < K; k++) acc[k] = fo[k];
                                any = true;
                            }
                        }
                    } else {
                        PosInfo<KM> pi;
                        std::memcpy(pi.pos, c1.pos, sizeof(double) * K);
                        std::memcpy(pi.vel, c1.vel, sizeof(double) * K);
                        pi.species = c1.species;
                        pi.cell_index = n;
                        pi.sender = v;
                        pi.receiver = w;
                        auto& box = inbox[ii][I.owner[w]];
                        std::lock_guard<std::mutex> lk(box.m);
                        box.pos.push_back(pi);
                    }
                }
                // The reference weighting drops `acc` (update_mechanics.rs:206-220).  The
                // physical weighting delivers it, as a cross-subdomain response would.
                if (any && cfg.weighting == CRM_WEIGHT_PHYSICAL)
                    for (int k = 0; k < K; k++) c1.force[k] = c1.force[k] + acc[k];
            }
        }
        // calculate_custom_domain_force runs in step 1 after the subdomain's own
        // pairs (run_sim.rs:277-279, update_mechanics.rs:244-274): aux.add_force(f)
        if (I.dforce.kind != CRM_DOMAIN_FORCE_NONE)
            for (uint64_t v : I.sub_voxels[sub])
                for (auto& c : I.voxels[v]) {
                    double f[KM] = {};
                    domain_force(I, c, f);
                    for (int k = 0; k < K; k++) c.force[k] = c.force[k] + f[k];
                }
        return 0;
    }

    // SubDomainForce::calculate_custom_force per point / rod vertex.
    //   spring: the test domain of cellular_raza/tests/domain_force_spring.rs:56-67,
    //           generalised per axis: (-strength_k) * (x_k - center_k);
    //   rod gel: CartesianSubDomainRods (bacterial_rods.rs:447-477): the row starts
    //           at (0, 0, -gel_pressure); when |v| != 0 and |p_k - min_k| <= d or
    //           |p_k - max_k| <= d on some axis, it gets += (-(v / |v|) * gel) * friction.
    void domain_force(const Instance& I, const Cell<KM>& c, double* f) const {
        const crm_domain_force& F = I.dforce;
        for (int v = 0; v < nv; v++) {
            const double* p = c.pos + v * D;
            double* o = f + v * D;
            if (F.kind == CRM_DOMAIN_FORCE_SPRING) {
                for (int k = 0; k < D; k++) o[k] = (-F.strength[k]) * (p[k] - F.center[k]);
                continue;
            }
            const double* vel = c.vel + v * D;
            for (int k = 0; k < D; k++) o[k] = 0.0;
            o[D - 1] = -F.gel_pressure;
            bool near = false;
            for (int k = 0; k < D; k++) {
                near = near || std::fabs(p[k] - I.grid.min[k]) <= F.surface_friction_distance;
                near = near || std::fabs(p[k] - I.grid.max[k]) <= F.surface_friction_distance;
            }
            double q2 = 0.0;
            for (int k = 0; k < D; k++) q2 = q2 + vel[k] * vel[k];
            const double q = std::sqrt(q2);
            if (q != 0.0 && near)
                for (int k = 0; k < D; k++) o[k] = o[k] + ((-(vel[k] / q)) * F.gel_pressure) * F.surface_friction;
        }
    }

    int set_domain_force(uint32_t ii, const crm_domain_force& f) override {
        if (ii >= inst.size() || f.kind > CRM_DOMAIN_FORCE_ROD_GEL) return CRM_INVALID_ARGUMENT;
        if (f.kind == CRM_DOMAIN_FORCE_ROD_GEL && !(nv > 1 && D == 3)) return CRM_INVALID_ARGUMENT;
        inst[ii].dforce = f;
        return CRM_OK;
    }

    // update_mechanics_interaction_step_2 (update_mechanics.rs:285-340)
    int step2(size_t ii, uint32_t sub) {
        Instance& I = inst[ii];
        auto& box = inbox[ii][sub];
        std::vector<PosInfo<KM>> recv;
        {
            std::lock_guard<std::mutex> lk(box.m);
            recv.swap(box.pos);
        }
        std::stable_sort(recv.begin(), recv.end(),
                         [](const PosInfo<KM>& a, const PosInfo<KM>& b) { return a.sender < b.sender; });
        for (auto& pi : recv) {
            ForceInfo<KM> fi;
            bool has;
            if (force_external(I, I.voxels[pi.receiver], pi.pos, pi.species, fi.force, has)) return 1;
            if (has) {
                fi.cell_index = pi.cell_index;
                fi.sender = pi.sender;
                auto& dst = inbox[ii][I.owner[pi.sender]];
                std::lock_guard<std::mutex> lk(dst.m);
                dst.force.push_back(fi);
            }
        }
        return 0;
    }

    // update_mechanics_interaction_step_3 (update_mechanics.rs:344-382)
    void step3(size_t ii, uint32_t sub) {
        Instance& I = inst[ii];
        auto& box = inbox[ii][sub];
        std::vector<ForceInfo<KM>> recv;
        {
            std::lock_guard<std::mutex> lk(box.m);
            recv.swap(box.force);
        }
        std::stable_sort(recv.begin(), recv.end(),
                         [](const ForceInfo<KM>& a, const ForceInfo<KM>& b) { return a.sender < b.sender; });
        for (auto& fi : recv) {
            Cell<KM>& c = I.voxels[fi.sender][fi.cell_index];
            for (int k = 0; k < K; k++) c.force[k] = c.force[k] + fi.force[k];
        }
    }

    // Mechanics::calculate_increment + get_random_contribution.
    void increment(const Instance& I, const Cell<KM>& c, const double* force, double* dx, double* dv,
                   double* dx_rand, double* dv_rand) const {
        const crm_mechanics& M = I.mech;
        const double dt = cfg.dt;
        double z[KM];
        bool have_noise = false;
        if (cfg.noise_mode == CRM_NOISE_BUFFER && noise) {
            for (int k = 0; k < K; k++) z[k] = noise[c.id * K + k];
            have_noise = true;
        } else if (cfg.noise_mode == CRM_NOISE_PHILOX) {
            gaussians(cfg.seed, c.id, step, K, z);
            have_noise = true;
        }
        for (int k = 0; k < K; k++) dx_rand[k] = dv_rand[k] = 0.0;
        switch (M.kind) {
            case CRM_MECHANICS_NEWTON_DAMPED:  // mechanics.rs:158-173
                for (int k = 0; k < K; k++) {
                    dx[k] = c.vel[k];
                    dv[k] = force[k] / M.mass - M.damping * c.vel[k];
                }
                break;
            case CRM_MECHANICS_LANGEVIN: {  // mechanics.rs:434-463
                double inv_m = 1.0 / M.mass;
                for (int k = 0; k < K; k++) {
                    dx[k] = c.vel[k];
                    double dv1 = inv_m * force[k];
                    double dv2 = (-M.damping) * c.vel[k];
                    dv[k] = dv1 + dv2;
                }
                if (have_noise) {
                    double coeff = std::sqrt(2.0 * M.damping * M.kb_temperature / M.mass);
                    double sd = std::sqrt(dt);
                    for (int k = 0; k < K; k++) dv_rand[k] = coeff * ((0.0 + sd * z[k]) / dt);
                }
                break;
            }
            case CRM_MECHANICS_BROWNIAN: {  // mechanics.rs:336-357
                double q = M.diffusion_constant / M.kb_temperature;
                for (int k = 0; k < K; k++) {
                    dx[k] = q * force[k];
                    dv[k] = 0.0;
                }
                if (have_noise) {
                    double coeff = std::sqrt(2.0 * M.diffusion_constant);
                    double sd = std::sqrt(dt);
                    for (int k = 0; k < K; k++) dx_rand[k] = coeff * ((0.0 + sd * z[k]) / dt);
                }
                break;
            }
            case CRM_MECHANICS_ROD: {  // bacterial_rods.rs:157-240
                crm_mechanics Mc = M;
                Mc.spring_length = c.spring_length;
                rod_increment(D, nv, Mc, c.pos, c.vel, force, dx, dv);
                if (have_noise) {
                    double sd = std::sqrt(dt);
                    for (int k = 0; k < K; k++)
                        dx_rand[k] = (((0.0 + sd * z[k]) * 0.25) * M.diffusion_constant) / dt;
                }
                break;
            }
        }
    }

    // chili solvers (solvers.rs:23-55,156-347) via local_mechanics_update (update_mechanics.rs:530-559)
    void integrate(const Instance& I, Cell<KM>& c) {
        const double dt = cfg.dt;
        double force[KM];
        std::memcpy(force, c.force, sizeof(double) * K);
        for (int k = 0; k < K; k++) c.force[k] = 0.0;  // get_current_force_and_reset
        double dx[KM], dv[KM], dxr[KM], dvr[KM];
        increment(I, c, force, dx, dv, dxr, dvr);
        ring_push(c, dx, dv);
        int branch;  // 0 euler, 1 ab2, 2 ab3
        if (cap == 0) branch = 0;
        else if (cap == 1) branch = (c.ring_size == 1) ? 1 : 0;
        else branch = (c.ring_size == 2) ? 2 : ((c.ring_size == 1) ? 1 : 0);
        const double f0 = 23.0 / 12.0, f1 = (-16.0) / 12.0, f2 = 5.0 / 12.0;
        const double g0 = 3.0 / 2.0, g1 = (-1.0) / 2.0;
        for (int k = 0; k < K; k++) {
            double x = c.pos[k], v = c.vel[k];
            double nx, nvv;
            if (branch == 0) {
                nx = (dx[k] * dt + x) * 1.0 + dxr[k] * dt;
                nvv = (dv[k] * dt + v) * 1.0 + dvr[k] * dt;
            } else if (branch == 1) {
                int s0 = ring_slot(c, 0);
                double h = c.ring_dx[s0][k], hv = c.ring_dv[s0][k];
                nx = ((dx[k] * g0 + h * g1) * dt + x) * 1.0 + dxr[k] * dt;
                nvv = ((dv[k] * g0 + hv * g1) * dt + v) * 1.0 + dvr[k] * dt;
            } else {
                int s0 = ring_slot(c, 0), s1 = ring_slot(c, 1);
                double h1 = c.ring_dx[s0][k], h2 = c.ring_dx[s1][k];
                double hv1 = c.ring_dv[s0][k], hv2 = c.ring_dv[s1][k];
                nx = (dx[k] * f0 + (h1 * f1 + h2 * f2)) * dt + x;
                nx = nx * 1.0 + dxr[k] * dt;
                nvv = (dv[k] * f0 + (hv1 * f1 + hv2 * f2)) * dt + v;
                nvv = nvv * 1.0 + dvr[k] * dt;
            }
            c.pos[k] = nx;
            if (I.mech.kind != CRM_MECHANICS_BROWNIAN) c.vel[k] = nvv;  // Brownian ignores set_velocity
        }
    }

    // CartesianSubDomain::apply_boundary (cartesian_cuboid_n.rs:594-631), per
    // vertex for rods (bacterial_rods.rs:489-533).  Returns 1 on BoundaryError.
    int apply_boundary(const Instance& I, Cell<KM>& c) const {
        const Grid& g = I.grid;
        for (int v = 0; v < nv; v++) {
            double* p = c.pos + v * D;
            double* u = c.vel + v * D;
            for (int i = 0; i < D; i++) {
                if (p[i] < g.min[i]) {
                    p[i] = 2.0 * g.min[i] - p[i];
                    u[i] = std::fabs(u[i]);
                }
                if (p[i] > g.max[i]) {
                    p[i] = 2.0 * g.max[i] - p[i];
                    u[i] = -std::fabs(u[i]);
                }
            }
        }
        for (int v = 0; v < nv; v++) {
            const double* p = c.pos + v * D;
            for (int i = 0; i < D; i++)
                if (p[i] < g.min[i] || p[i] > g.max[i]) return 1;
        }
        return 0;
    }

    // Point used for binning: the position, or the vertex mean of a rod
    // (bacterial_rods.rs:368-372,550-554: row_sum / nrows).
    void bin_point(const Cell<KM>& c, double* p) const {
        if (nv == 1) {
            for (int i = 0; i < D; i++) p[i] = c.pos[i];
            return;
        }
        for (int i = 0; i < D; i++) {
            double s = 0.0;
            for (int v = 0; v < nv; v++) s = s + c.pos[v * D + i];
            p[i] = s / (double)nv;
        }
    }

    // sort_cells_in_voxels_step_1 (update_mechanics.rs:418-477).  Partition errors
    // go into a vector that is never read (update_mechanics.rs:430), so a cell
    // whose voxel cannot be resolved stays where it is.
    void sort_step1(size_t ii, uint32_t sub) {
        Instance& I = inst[ii];
        std::vector<std::pair<uint64_t, Cell<KM>>> movers;
        double p[3];
        uint64_t idx[3];
        for (uint64_t v : I.sub_voxels[sub]) {
            auto& cells = I.voxels[v];
            size_t keep = 0;
            for (size_t j = 0; j < cells.size(); j++) {
                bin_point(cells[j], p);
                bool move = false;
                if (index_of(I.grid, p, idx) == 0 && in_grid(I.grid, idx)) move = plain_of(I.grid, idx) != v;
                if (move) movers.emplace_back(v, cells[j]);
                else {
                    if (keep != j) cells[keep] = cells[j];
                    keep++;
                }
            }
            cells.resize(keep);
        }
        for (auto& mv : movers) {
            bin_point(mv.second, p);
            index_of(I.grid, p, idx);
            uint64_t w = plain_of(I.grid, idx);
            if (I.owner[w] == sub) {
                I.voxels[w].push_back(mv.second);
            } else {
                SendCell<KM> sc;
                sc.sender = mv.first;
                sc.cell = mv.second;
                auto& box = inbox[ii][I.owner[w]];
                std::lock_guard<std::mutex> lk(box.m);
                box.cells.push_back(sc);
            }
        }
    }

    // sort_cells_in_voxels_step_2 (update_mechanics.rs:485-513)
    void sort_step2(size_t ii, uint32_t sub) {
        Instance& I = inst[ii];
        auto& box = inbox[ii][sub];
        std::vector<SendCell<KM>> recv;
        {
            std::lock_guard<std::mutex> lk(box.m);
            recv.swap(box.cells);
        }
        std::stable_sort(recv.begin(), recv.end(),
                         [](const SendCell<KM>& a, const SendCell<KM>& b) { return a.sender < b.sender; });
        double p[3];
        uint64_t idx[3];
        for (auto& sc : recv) {
            bin_point(sc.cell, p);
            index_of(I.grid, p, idx);
            I.voxels[plain_of(I.grid, idx)].push_back(sc.cell);
        }
    }

    // One subdomain's share of one step (run_sim.rs:380-403).  Errors of the force
    // phase and of the integrate/boundary phase go to separate words, each read
    // only after the barrier that closes its phase, so every thread sees the same
    // value (the reference's BarrierSync error flag, simulation_flow.rs:397-425).
    void run_sub(size_t ii, uint32_t sub, Barrier* bar, StepError& Ef, StepError& Em, bool forces_only) {
        Instance& I = inst[ii];
        auto sync = [&]() {
            if (bar) bar->wait();
        };
        if (step1(ii, sub)) Ef.set(CRM_CALC_ERROR, "CalcError: identical position for two objects");
        sync();
        if (step2(ii, sub)) Ef.set(CRM_CALC_ERROR, "CalcError: identical position for two objects");
        sync();
        step3(ii, sub);
        if (forces_only) return;
        sync();
        if (Ef.code.load()) return;
        for (uint64_t v : I.sub_voxels[sub])
            for (auto& c : I.voxels[v]) integrate(I, c);
        if (cycle_on) {
            // local funcs after the mechanics update: react_to_neighbors, then the
            // cycle update (run_sim.rs:259-283); then update_cell_cycle_4 (step 4)
            for (uint64_t v : I.sub_voxels[sub])
                for (auto& c : I.voxels[v]) react_and_grow(c);
            for (uint64_t v : I.sub_voxels[sub])
                if (cycle_voxel(I, v)) Em.set(CRM_DIVISION_ERROR, "DivisionError: Could not calculate new segment lengths");
        }
        for (uint64_t v : I.sub_voxels[sub])
            for (auto& c : I.voxels[v])
                if (apply_boundary(I, c)) Em.set(CRM_BOUNDARY_ERROR, "BoundaryError: Particle is out of domain");
        sync();
        if (Em.code.load()) return;
        if (cycle_on) {
            if (sub == 0) number_new_cells(I);
            sync();
        }
        sort_step1(ii, sub);
        sync();
        sort_step2(ii, sub);
    }

    // The accumulator of the rods example: (own vertex, other vertex) pairs of the
    // stencil closer than the sensing radius, on the positions the forces see.
    void sense_all() {
        std::vector<uint32_t> cnt(n_agents);
        neighbor_counts(cyc.sensing_radius, cnt.data());
        for_each_cell([&](Cell<KM>& c) { c.sensed = cnt[c.id]; });
    }

    // react_to_neighbors then Cycle::update_cycle (bacterial_rods/src/main.rs:66-73,
    // 88-100): growth = ((base * (sat - n)) / sat).max(0) when n > 0, else base;
    // spring_length += growth * dt; Division when spring_length > threshold.
    void react_and_grow(Cell<KM>& c) {
        if (c.sensed > 0) {
            const double g = (cyc.growth_base_rate * (cyc.neighbor_saturation - (double)c.sensed)) /
                             cyc.neighbor_saturation;
            c.growth_rate = std::fmax(g, 0.0);
        } else {
            c.growth_rate = cyc.growth_base_rate;
        }
        c.sensed = 0;  // clear_accumulator
        c.spring_length = c.spring_length + c.growth_rate * cfg.dt;
        if (c.spring_length > cyc.division_threshold) c.divide = true;
    }

    // RodMechanics::divide (bacterial_rods.rs:586-662) on c (the mother, changed in
    // place) into d (the daughter, otherwise a clone).  Returns 1 on DivisionError.
    int rod_divide(Cell<KM>& c, double radius, Cell<KM>& d) const {
        const int n = nv;
        double pos[KM];
        std::memcpy(pos, c.pos, sizeof(double) * K);
        d = c;
        const double two = 1.0 + 1.0, one_half = 1.0 / two;
        const double div_factor = one_half - radius / ((double)n * c.spring_length);
        const double nsl = div_factor * c.spring_length;
        c.spring_length = nsl;
        d.spring_length = nsl;
        double seg[CRM_MAX_ROD_VERTICES] = {};
        for (int i = 0; i + 1 < n; i++) {
            double z[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < D; k++) z[k] = pos[i * D + k] - pos[(i + 1) * D + k];
            double q2 = 0.0;
            for (int k = 0; k < D; k++) q2 = q2 + z[k] * z[k];
            seg[i] = std::sqrt(q2);
        }
        auto head = [&](int m) {  // segments.iter().take(m).sum()
            double t = 0.0;
            for (int i = 0; i < m; i++) t = t + seg[i];
            return t;
        };
        auto tail = [&](int m) {  // segments.iter().rev().take(m).sum()
            double t = 0.0;
            for (int i = 0; i < m; i++) t = t + seg[n - 2 - i];
            return t;
        };
        const double seg_len = (head(n - 1) - two * radius) / (double)(n - 1) / two;
        for (int v = 0; v < n; v++) {
            const double L = (double)v * seg_len;
            int k = -1;
            for (int m = 0; m < n - 1; m++)
                if (head(m) <= L) k = m;
            if (k < 0) return 1;
            const double q = (L - head(k)) / seg[k];
            for (int j = 0; j < D; j++) c.pos[v * D + j] = pos[k * D + j] * (1.0 - q) + pos[(k + 1) * D + j] * q;
            int m = -1;
            for (int mm = 0; mm < n; mm++)
                if (tail(mm) <= L) m = mm;
            if (m < 0 || n - m - 2 < 0) return 1;
            const double p = (L - tail(m)) / seg[n - m - 2];
            for (int j = 0; j < D; j++)
                d.pos[(n - v - 1) * D + j] = pos[(n - m - 1) * D + j] * (1.0 - p) + pos[(n - m - 2) * D + j] * p;
        }
        return 0;
    }

    // Voxel::update_cell_cycle_4 (update_cycle.rs:10-69): mothers in voxel order get
    // Division(voxel, ++counter) and their old identifier as parent; the daughters
    // are appended after the loop with Division(voxel, ++counter), the same parent
    // and a fresh AuxStorage (empty history, zero force).
    int cycle_voxel(Instance& I, uint64_t v) {
        auto& cells = I.voxels[v];
        std::vector<Cell<KM>> born;
        for (auto& c : cells) {
            if (!c.divide) continue;
            c.divide = false;
            Cell<KM> d;
            if (rod_divide(c, cyc.division_radius, d)) return 1;
            uint64_t par[3] = {c.ident[0], c.ident[1], c.ident[2]};
            I.id_counter[v] += 1;
            c.ident[0] = 1;
            c.ident[1] = v;
            c.ident[2] = I.id_counter[v];
            std::memcpy(c.parent, par, sizeof(par));
            std::memcpy(d.parent, par, sizeof(par));
            d.divide = false;
            d.sensed = 0;
            d.ring_size = 0;
            d.ring_head = 0;
            for (int k = 0; k < K; k++) d.force[k] = 0.0;
            d.id = UINT64_MAX;  // numbered after every voxel has divided
            born.push_back(d);
        }
        for (auto& d : born) {
            I.id_counter[v] += 1;
            d.ident[0] = 1;
            d.ident[1] = v;
            d.ident[2] = I.id_counter[v];
            cells.push_back(d);
        }
        return 0;
    }

    // Agent indices of the cells born this step: n_agents, n_agents + 1, ... in
    // voxel order, then order inside the voxel.
    void number_new_cells(Instance& I) {
        for (auto& cells : I.voxels)
            for (auto& c : cells)
                if (c.id == UINT64_MAX) c.id = n_agents++;
    }

    int set_rod_cycle(const crm_rod_cycle& y) override {
        if (nv < 2) return CRM_INVALID_ARGUMENT;
        cyc = y;
        cycle_on = true;
        for_each_cell([&](Cell<KM>& c) { c.growth_rate = y.growth_base_rate; });
        return CRM_OK;
    }

    int set_spring_lengths(const double* sl) override {
        for_each_cell([&](Cell<KM>& c) { c.spring_length = sl[c.id]; });
        return CRM_OK;
    }

    // CycleEvent::Remove as update_cell_cycle_4 applies it (update_cycle.rs:41-55):
    // the flagged cells leave their voxel's Vec, retain keeps the order of the
    // others; identifiers and id counters stay.  Agent indices are compacted in
    // order (agent a becomes the number of survivors before a).
    int remove_agents(const uint8_t* flags) override {
        std::vector<uint64_t> before(n_agents);
        uint64_t kept = 0;
        for (uint64_t a = 0; a < n_agents; a++) {
            before[a] = kept;
            kept += flags[a] ? 0 : 1;
        }
        if (kept == n_agents) return CRM_OK;
        if (kept == 0) {
            err = "the removal would leave no agents";
            return CRM_INVALID_ARGUMENT;
        }
        for (auto& I : inst)
            for (auto& cells : I.voxels) {
                std::vector<Cell<KM>> keep;
                keep.reserve(cells.size());
                for (const Cell<KM>& c : cells)
                    if (!flags[c.id]) keep.push_back(c);
                for (Cell<KM>& c : keep) c.id = before[c.id];
                cells.swap(keep);
            }
        n_agents = kept;
        return CRM_OK;
    }

    // CycleEvent::Division of a caller's Cycle for point agents, as
    // Voxel::update_cell_cycle_4 applies it (update_cycle.rs:28-66), between steps:
    // voxel by voxel in list order each mother takes the state C::divide left it and
    // Division(voxel, ++id_counter) with its old identifier as parent; the voxel's
    // daughters follow the loop in the mothers' order with Division(voxel,
    // ++id_counter), the same parent and a fresh AuxStorage (empty ring).  Then
    // apply_boundary and sort_cells_in_voxels (update_mechanics.rs:418-513) as at
    // the end of a step.  Daughter i is agent n_agents + i.
    int insert_agents(uint64_t n, const uint64_t* mother, const double* mpos, const double* mvel,
                      const double* dpos, const double* dvel, const int32_t* dsp) override {
        if (nv != 1) return CRM_INVALID_ARGUMENT;
        std::vector<int64_t> call(n_agents, -1);
        for (uint64_t i = 0; i < n; i++) {
            if (mother[i] >= n_agents || call[mother[i]] >= 0) return CRM_INVALID_ARGUMENT;
            call[mother[i]] = (int64_t)i;
        }
        // the divided states first: on a BoundaryError nothing changes
        std::vector<Cell<KM>> mnew(n), dnew(n);
        for (auto& I : inst)
            for (auto& cells : I.voxels)
                for (auto& c : cells) {
                    if (call[c.id] < 0) continue;
                    const uint64_t i = (uint64_t)call[c.id];
                    Cell<KM> m = c, d;
                    std::memset(&d, 0, sizeof(d));
                    for (int k = 0; k < K; k++) {
                        if (mpos) m.pos[k] = mpos[i * K + k];
                        if (mvel) m.vel[k] = mvel[i * K + k];
                    }
                    d.id = n_agents + i;
                    d.species = dsp ? dsp[i] : c.species;
                    for (int k = 0; k < K; k++) {
                        d.pos[k] = dpos[i * K + k];
                        d.vel[k] = dvel ? dvel[i * K + k] : m.vel[k];
                    }
                    if (apply_boundary(I, m) || apply_boundary(I, d)) {
                        err = "BoundaryError: Particle is out of domain";
                        return CRM_BOUNDARY_ERROR;
                    }
                    mnew[i] = m;
                    dnew[i] = d;
                }
        for (auto& I : inst)
            for (size_t v = 0; v < I.voxels.size(); v++) {
                auto& cells = I.voxels[v];
                std::vector<Cell<KM>> born;
                for (auto& c : cells) {
                    if (call[c.id] < 0) continue;
                    const uint64_t i = (uint64_t)call[c.id];
                    const uint64_t par[3] = {c.ident[0], c.ident[1], c.ident[2]};
                    c = mnew[i];
                    I.id_counter[v] += 1;
                    c.ident[0] = 1;
                    c.ident[1] = v;
                    c.ident[2] = I.id_counter[v];
                    std::memcpy(c.parent, par, sizeof(par));
                    Cell<KM> d = dnew[i];
                    std::memcpy(d.parent, par, sizeof(par));
                    born.push_back(d);
                }
                for (auto& d : born) {
                    I.id_counter[v] += 1;
                    d.ident[0] = 1;
                    d.ident[1] = v;
                    d.ident[2] = I.id_counter[v];
                    cells.push_back(d);
                }
            }
        n_agents += n;
        for (size_t ii = 0; ii < inst.size(); ii++) {
            for (uint32_t sub = 0; sub < n_sub; sub++) sort_step1(ii, sub);
            for (uint32_t sub = 0; sub < n_sub; sub++) sort_step2(ii, sub);
        }
        return CRM_OK;
    }

    void cycle_out(double* spring_length, double* growth_rate, uint64_t* ident, uint64_t* parent) override {
        for_each_cell([&](Cell<KM>& c) {
            if (spring_length) spring_length[c.id] = c.spring_length;
            if (growth_rate) growth_rate[c.id] = c.growth_rate;
            for (int j = 0; j < 3; j++) {
                if (ident) ident[c.id * 3 + j] = c.ident[j];
                if (parent) parent[c.id * 3 + j] = c.parent[j];
            }
        });
    }

    int run_phase(bool forces_only) override {
        if (cycle_on && !forces_only) sense_all();
        for (size_t ii = 0; ii < inst.size(); ii++) {
            StepError Ef, Em;
            if (n_sub == 1) {
                run_sub(ii, 0, nullptr, Ef, Em, forces_only);
            } else {
                Barrier bar((int)n_sub);
                std::vector<std::thread> th;
                for (uint32_t s = 0; s < n_sub; s++)
                    th.emplace_back([&, s]() { run_sub(ii, s, &bar, Ef, Em, forces_only); });
                for (auto& t : th) t.join();
            }
            StepError& E = Ef.code.load() ? Ef : Em;
            if (E.code.load()) {
                err = E.msg;
                return E.code.load();
            }
        }
        return 0;
    }

    template <class F>
    void for_each_cell(F f) {
        for (auto& I : inst)
            for (auto& cells : I.voxels)
                for (auto& c : cells) f(c);
    }

    int init(const crm_instance* instances, uint32_t n_instances) override {
        inst.resize(n_instances);
        inbox.resize(n_instances);
        for (uint32_t i = 0; i < n_instances; i++) {
            Instance& I = inst[i];
            int st = make_grid(instances[i].domain, D, I.grid, err);
            if (st) return st;
            I.pot = instances[i].potential;
            I.mech = instances[i].mechanics;
            I.voxels.assign(I.grid.V, {});
            I.owner.assign(I.grid.V, 0);
            // create_subdomains: lexicographic chunks (cartesian_cuboid_n.rs:493-584)
            uint64_t rn, rm, avg;
            uint32_t nsub = n_sub;
            if (I.grid.V < nsub) {
                err = "DecomposeError: more subdomains than voxels";
                return CRM_SETUP_ERROR;
            }
            if (!decomp_res(I.grid.V, nsub, rn, rm, avg)) {
                err = "DecomposeError: Could not find a suiting decomposition";
                return CRM_SETUP_ERROR;
            }
            I.sub_voxels.assign(n_sub, {});
            uint64_t switcher = rn * avg;
            for (uint64_t p = 0; p < I.grid.V; p++) {
                uint64_t s = p < switcher ? p / avg : (p - switcher) / std::max<uint64_t>(avg - 1, 1) + rn;
                I.owner[p] = (uint32_t)s;
                I.sub_voxels[s].push_back(p);
            }
            inbox[i] = std::vector<Inbox>(n_sub);
        }
        return 0;
    }

    int set_agents(const uint64_t* counts, const double* pos, const double* vel,
                   const int32_t* species) override {
        uint64_t a = 0;
        for (auto& I : inst)
            for (auto& v : I.voxels) v.clear();
        for (size_t ii = 0; ii < inst.size(); ii++) {
            Instance& I = inst[ii];
            // Domain::decompose routes cells in input order; insert_initial_cells
            // appends them to their voxel (domain.rs:357-380, datastructures.rs:278-299).
            for (uint64_t j = 0; j < counts[ii]; j++, a++) {
                Cell<KM> c;
                std::memset(&c, 0, sizeof(c));
                c.id = a;
                c.species = species ? species[a] : 0;
                c.spring_length = I.mech.spring_length;
                c.growth_rate = cyc.growth_base_rate;
                c.ident[0] = 0;
                c.ident[1] = j;  // CellBox::new_initial(n_cell): index in this simulation's agents
                c.parent[0] = 3;
                for (int k = 0; k < K; k++) {
                    c.pos[k] = pos[a * K + k];
                    c.vel[k] = vel ? vel[a * K + k] : 0.0;
                }
                double p[3];
                uint64_t idx[3] = {0, 0, 0};
                bin_point(c, p);
                for (int i = 0; i < D; i++)
                    if (I.grid.min[i] >= p[i]) {
                        err = "BoundaryError: Min must be smaller than Max (agent " + std::to_string(a) + ")";
                        return CRM_BOUNDARY_ERROR;
                    }
                if (index_of(I.grid, p, idx)) {
                    err = "BoundaryError: Cannot convert float to usize (agent " + std::to_string(a) + ")";
                    return CRM_BOUNDARY_ERROR;
                }
                if (!in_grid(I.grid, idx)) {
                    err = "DecomposeError: could not find voxel index (agent " + std::to_string(a) + ")";
                    return CRM_INDEX_ERROR;
                }
                I.voxels[plain_of(I.grid, idx)].push_back(c);
            }
            // insert_initial_cells counts the initial cells of each voxel (datastructures.rs:286-297)
            I.id_counter.assign(I.voxels.size(), 0);
            for (size_t v = 0; v < I.voxels.size(); v++) I.id_counter[v] = I.voxels[v].size();
        }
        n_agents = a;
        step = 0;
        have_agents = true;
        return 0;
    }

    void zero_forces() override {
        for_each_cell([&](Cell<KM>& c) {
            for (int k = 0; k < K; k++) c.force[k] = 0.0;
        });
    }

    void forces_out(double* force) override {
        for_each_cell([&](Cell<KM>& c) {
            for (int k = 0; k < K; k++) force[c.id * K + k] = c.force[k];
        });
    }

    void state_out(double* pos, double* vel) override {
        for_each_cell([&](Cell<KM>& c) {
            for (int k = 0; k < K; k++) {
                if (pos) pos[c.id * K + k] = c.pos[k];
                if (vel) vel[c.id * K + k] = c.vel[k];
            }
        });
    }

    uint32_t history_out(double* dx, double* dv) override {
        uint32_t ns = 0;
        for_each_cell([&](Cell<KM>& c) {
            ns = (uint32_t)c.ring_size;
            for (int s = 0; s < 2; s++)
                for (int k = 0; k < K; k++) {
                    double a = 0.0, b = 0.0;
                    if (s < c.ring_size) {
                        int slot = (c.ring_head + s) % cap;  // oldest first
                        a = c.ring_dx[slot][k];
                        b = c.ring_dv[slot][k];
                    }
                    if (dx) dx[((uint64_t)s * n_agents + c.id) * K + k] = a;
                    if (dv) dv[((uint64_t)s * n_agents + c.id) * K + k] = b;
                }
        });
        return ns;
    }

    // NeighborSensing as in the rods example (bacterial_rods/src/main.rs:48-78): every
    // cell senses every other cell of its voxel and of the neighbouring voxels once
    // per step (update_mechanics.rs:199-202 same voxel, 105 adjacent voxels), and the
    // accumulator counts the (own point, ext point) pairs -- points, or rod
    // vertices -- with norm(own - ext) < radius.
    void neighbor_counts(double radius, uint32_t* counts) override {
        std::vector<uint64_t> nb;
        for (auto& I : inst) {
            for (uint64_t v = 0; v < I.voxels.size(); v++) {
                neighbors_of(I.grid, v, nb);
                nb.push_back(v);
                for (const auto& c : I.voxels[v]) {
                    uint32_t n = 0;
                    for (uint64_t w : nb)
                        for (const auto& e : I.voxels[w]) {
                            if (&e == &c) continue;
                            for (int a = 0; a < nv; a++)
                                for (int b = 0; b < nv; b++) {
                                    double d2 = 0.0;
                                    for (int k = 0; k < D; k++) {
                                        const double d = c.pos[a * D + k] - e.pos[b * D + k];
                                        d2 = d2 + d * d;
                                    }
                                    if (std::sqrt(d2) < radius) n++;
                                }
                        }
                    counts[c.id] = n;
                }
            }
        }
    }

    void voxels_out(uint32_t* voxel_of_agent, uint32_t* start_out, uint64_t* order_out) override {
        uint64_t off = 0, pos = 0;
        for (auto& I : inst) {
            for (uint64_t v = 0; v < I.grid.V; v++) {
                if (start_out) start_out[off + v] = (uint32_t)pos;
                for (auto& c : I.voxels[v]) {
                    if (voxel_of_agent) voxel_of_agent[c.id] = (uint32_t)(off + v);
                    if (order_out) order_out[pos] = c.id;
                    pos++;
                }
            }
            off += I.grid.V;
        }
        if (start_out) start_out[off] = (uint32_t)pos;
    }

    uint64_t n_voxels_total() override {
        uint64_t t = 0;
        for (auto& I : inst) t += I.grid.V;
        return t;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

int orc_grid(const crm_domain* d, uint32_t dim, uint64_t n_out[3], double dx_out[3]) {
    Grid g;
    std::string err;
    int st = make_grid(*d, (int)dim, g, err);
    if (st) return st;
    for (int i = 0; i < 3; i++) {
        n_out[i] = g.n[i];
        dx_out[i] = g.dx[i];
    }
    return 0;
}

// Voxel index of a point.  initial != 0 applies the decompose-time rules
// (strict min < p, index must lie in the grid); otherwise the per-step rule
// (conversion only).  Returns 0 ok, CRM_BOUNDARY_ERROR, CRM_INDEX_ERROR.
int orc_voxel_index(const crm_domain* d, uint32_t dim, const double* p, int initial, uint64_t idx_out[3]) {
    Grid g;
    std::string err;
    int st = make_grid(*d, (int)dim, g, err);
    if (st) return st;
    if (initial)
        for (uint32_t i = 0; i < dim; i++)
            if (g.min[i] >= p[i]) return CRM_BOUNDARY_ERROR;
    for (int i = 0; i < 3; i++) idx_out[i] = 0;
    if (index_of(g, p, idx_out)) return CRM_BOUNDARY_ERROR;
    if (initial && !in_grid(g, idx_out)) return CRM_INDEX_ERROR;
    return 0;
}

uint32_t orc_neighbors(const uint64_t n[3], uint32_t dim, uint64_t plain, uint64_t* out) {
    Grid g;
    g.D = (int)dim;
    g.V = 1;
    for (int i = 0; i < 3; i++) g.n[i] = i < (int)dim ? n[i] : 1;
    std::vector<uint64_t> nb;
    neighbors_of(g, plain, nb);
    for (size_t i = 0; i < nb.size(); i++) out[i] = nb[i];
    return (uint32_t)nb.size();
}

int orc_decomp_res(uint64_t n_voxel, uint64_t n_regions, uint64_t out[3]) {
    return decomp_res(n_voxel, n_regions, out[0], out[1], out[2]) ? 0 : 1;
}

int orc_pair_force(const crm_potential* P, uint32_t dim, const double* own, const double* ext, int32_t sp_own,
                   int32_t sp_ext, double* f_own, double* f_ext) {
    return point_force((int)dim, *P, own, ext, sp_own, sp_ext, f_own, f_ext);
}

int orc_rod_pair_force(const crm_potential* P, uint32_t dim, uint32_t nv, const double* own, const double* ext,
                       double* f_own, double* f_ext) {
    return rod_force((int)dim, (int)nv, *P, own, ext, 0, 0, f_own, f_ext);
}

void orc_nearest_point(uint32_t dim, const double* p, const double* a, const double* b, double* dist,
                       double* x, double* t) {
    nearest_point((int)dim, p, a, b, *dist, x, *t);
}

void orc_rod_increment(const crm_mechanics* M, uint32_t dim, uint32_t nv, const double* pos, const double* vel,
                       const double* force, double* dx, double* dv) {
    rod_increment((int)dim, (int)nv, *M, pos, vel, force, dx, dv);
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr, key, out);
}

void orc_gaussians(uint64_t seed, uint64_t agent, uint64_t step, uint32_t n, double* out) {
    gaussians(seed, agent, step, (int)n, out);
}

int orc_create(const crm_config* cfg, const crm_instance* instances, uint32_t n_instances,
               uint32_t n_subdomains, void** out) {
    *out = nullptr;
    if (!cfg || !instances || n_instances == 0 || n_subdomains == 0) return CRM_INVALID_ARGUMENT;
    if (cfg->dim < 2 || cfg->dim > 3) return CRM_INVALID_ARGUMENT;
    if (cfg->solver_order > 2) return CRM_INVALID_ARGUMENT;
    bool rods = instances[0].mechanics.kind == CRM_MECHANICS_ROD;
    int nv = rods ? (int)instances[0].mechanics.n_vertices : 1;
    if (nv < 1 || nv > CRM_MAX_ROD_VERTICES || (rods && nv < 2)) return CRM_INVALID_ARGUMENT;
    SimBase* S;
    if (rods) S = new Sim<KMAX>();
    else S = new Sim<3>();
    S->cfg = *cfg;
    S->D = (int)cfg->dim;
    S->nv = nv;
    S->K = nv * S->D;
    S->cap = (int)cfg->solver_order;
    S->n_sub = n_subdomains;
    int st = S->init(instances, n_instances);
    if (st) {
        delete S;
        return st;
    }
    *out = S;
    return 0;
}

int orc_set_agents(void* h, const uint64_t* counts, const double* pos, const double* vel,
                   const int32_t* species) {
    SimBase& S = *(SimBase*)h;
    return S.set_agents(counts, pos, vel, species);
}

int orc_step(void* h, uint32_t n_steps, const double* noise) {
    SimBase& S = *(SimBase*)h;
    if (!S.have_agents) return CRM_INVALID_ARGUMENT;
    for (uint32_t s = 0; s < n_steps; s++) {
        S.noise = noise ? noise + (uint64_t)s * S.n_agents * S.K : nullptr;
        int st = S.run_phase(false);
        S.noise = nullptr;
        if (st) return st;
        S.step++;
    }
    return 0;
}

int orc_step_timed(void* h, uint32_t n_steps, double* seconds) {
    auto t0 = std::chrono::steady_clock::now();
    int st = orc_step(h, n_steps, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    return st;
}

int orc_compute_forces(void* h, double* force) {
    SimBase& S = *(SimBase*)h;
    if (!S.have_agents) return CRM_INVALID_ARGUMENT;
    S.zero_forces();
    int st = S.run_phase(true);
    S.forces_out(force);
    S.zero_forces();
    return st;
}

int orc_get_state(void* h, double* pos, double* vel) {
    ((SimBase*)h)->state_out(pos, vel);
    return 0;
}

int orc_get_history(void* h, double* dx, double* dv, uint32_t* n_stored) {
    uint32_t ns = ((SimBase*)h)->history_out(dx, dv);
    if (n_stored) *n_stored = ns;
    return 0;
}

// Voxel (instance offset + plain index) of every agent, and per voxel the agent
// indices in the voxel's Vec order (order_out, n_agents entries, CSR by start_out
// with n_voxels_total + 1 entries); any pointer may be NULL.
int orc_get_voxels(void* h, uint32_t* voxel_of_agent, uint32_t* start_out, uint64_t* order_out) {
    ((SimBase*)h)->voxels_out(voxel_of_agent, start_out, order_out);
    return 0;
}

uint64_t orc_n_voxels_total(void* h) { return ((SimBase*)h)->n_voxels_total(); }

// Neighbour count of every agent (by agent index) for the current positions.
int orc_set_rod_cycle(void* h, const crm_rod_cycle* y) {
    if (!h || !y) return CRM_INVALID_ARGUMENT;
    return static_cast<SimBase*>(h)->set_rod_cycle(*y);
}

int orc_set_spring_lengths(void* h, const double* sl) {
    if (!h || !sl) return CRM_INVALID_ARGUMENT;
    return static_cast<SimBase*>(h)->set_spring_lengths(sl);
}

int orc_remove_agents(void* h, const uint8_t* flags) {
    if (!h || !flags) return CRM_INVALID_ARGUMENT;
    return static_cast<SimBase*>(h)->remove_agents(flags);
}

int orc_insert_agents(void* h, uint64_t n, const uint64_t* mother, const double* mpos, const double* mvel,
                      const double* dpos, const double* dvel, const int32_t* dsp) {
    if (!h || (n && (!mother || !dpos))) return CRM_INVALID_ARGUMENT;
    return static_cast<SimBase*>(h)->insert_agents(n, mother, mpos, mvel, dpos, dvel, dsp);
}

int orc_cycle_out(void* h, double* spring_length, double* growth_rate, uint64_t* ident, uint64_t* parent) {
    if (!h) return CRM_INVALID_ARGUMENT;
    static_cast<SimBase*>(h)->cycle_out(spring_length, growth_rate, ident, parent);
    return CRM_OK;
}

int orc_set_domain_force(void* h, uint32_t inst, const crm_domain_force* f) {
    if (!h || !f) return CRM_INVALID_ARGUMENT;
    return static_cast<SimBase*>(h)->set_domain_force(inst, *f);
}

int orc_neighbor_counts(void* h, double radius, uint32_t* counts) {
    SimBase& S = *(SimBase*)h;
    if (!S.have_agents) return CRM_INVALID_ARGUMENT;
    S.neighbor_counts(radius, counts);
    return 0;
}

uint64_t orc_n_agents(void* h) { return ((SimBase*)h)->n_agents; }

size_t orc_last_error(void* h, char* buf, size_t len) {
    SimBase& S = *(SimBase*)h;
    if (buf && len) {
        size_t n = std::min(len - 1, S.err.size());
        std::memcpy(buf, S.err.data(), n);
        buf[n] = 0;
    }
    return S.err.size();
}

void orc_destroy(void* h) { delete (SimBase*)h; }

}  // extern "C"
