// params.hpp -- per-instance constants and the device control block.
//
// Plain structs copied once to device memory at crm_create.  Grids are stored
// as 3-D with a leading extent of 1 for 2-D domains, so plain indices are the
// same as the reference's lexicographic ones (datastructures.rs:113-120):
// plain = (i0 * n1 + i1) * n2 + i2, and a 2-D point (x, y) sits on axes 1, 2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crmech.h"

namespace crm {

template <typename T>
struct PotParams {
    T radius, two_radius, stiffness, cutoff, strength, bound;
    T en, em, mie_sigma, mie_constant;
    T mie_en_s, mie_em_s;  // en / mie_sigma, em / mie_sigma (pair-independent)
    T epsilon, sigma;
    T rel_range, sign;
    T species_matrix[CRM_MAX_SPECIES * CRM_MAX_SPECIES];
    int ien1, iem1;  // integer exponents en+1, em+1 when exact, else -1
    int pad0, pad1;
};

struct InstParams {
    // grid, 3-D layout (axis 0 has extent 1 in 2-D)
    double min[3], max[3], dx[3];
    double qinv[3];   // 2^30 / dx: fixed-point scale of voxel-relative positions
    double inv_dx[3]; // 1 / dx (voxel index estimate; see axis_index_fast)
    double inv_n2, inv_n12;  // 1 / n[2], 1 / (n[1] n[2]) (plain-index decomposition)
    float qunit[3];   // dx / 2^30
    float dxf[3];
    uint32_t n[3];
    uint32_t voxel_offset;  // first global voxel of this instance
    uint32_t n_voxels;
    uint64_t agent_offset;  // first agent slot of this instance (slots never change instance)
    uint64_t n_agents;
    PotParams<double> pd;
    PotParams<float> pf;
    // mechanics
    double mass, damping, kb_temperature, diffusion_constant;
    double spring_tension, rigidity, spring_length;
    double inv_mass;        // 1 / mass (Langevin, mechanics.rs:457)
    double langevin_coeff;  // sqrt(2 * damping * kT / mass) (mechanics.rs:439-445)
    double brownian_q;      // diffusion / kT (mechanics.rs:352)
    double brownian_coeff;  // sqrt(2 * diffusion) (mechanics.rs:341)
    // domain force (SubDomainForce), added after the pair forces
    uint32_t dforce;        // crm_domain_force_kind
    uint32_t dpad;
    double dk[3], dc[3];    // spring: strength and centre per axis, 3-D layout
    double gel, fric, fric_d;  // rods gel: pressure, surface friction, friction distance
    // chili subdomains (create_subdomains with get_decomp_res(V, k) = (n, m, avg)):
    // plain index p belongs to chunk p / avg below sub_switch = n * avg, else to
    // (p - sub_switch) / sub_avg1 + n with sub_avg1 = max(avg - 1, 1).  sub_on = 0:
    // one subdomain.
    uint32_t sub_on, sub_avg, sub_avg1, sub_switch, sub_n, sub_pad;
};

// Subdomain (chili thread) owning the voxel with plain index p of an instance.
__host__ __device__ __forceinline__ uint32_t subdomain_of(const InstParams& P, uint32_t p) {
    return p < P.sub_switch ? p / P.sub_avg : (p - P.sub_switch) / P.sub_avg1 + P.sub_n;
}

// Work unit of the agent-parallel kernels: a block of consecutive slots inside
// one instance, so per-instance parameters are wave-uniform.
struct BlockWork {
    uint32_t inst;
    uint32_t count;
    uint64_t first;
};

// Device control block.  `step` is the number of completed steps; the step
// kernels read it, so a captured graph replays correctly.
struct Control {
    uint64_t step;
    uint64_t error_key;     // min over (step << 32 | agent index) of failing agents
    uint32_t error_flags;   // bit (1 << crm_status) per error kind seen
    uint32_t n_div;         // rods dividing in the current step (cycle)
    uint64_t out_of_grid;   // agents that kept their voxel because the new index had no voxel
    uint32_t skip;          // set by k_scan_tiles: this step's sort is skipped (an earlier step failed)
    uint32_t pad;
};

}  // namespace crm
