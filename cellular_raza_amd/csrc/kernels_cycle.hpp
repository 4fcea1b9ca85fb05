// kernels_cycle.hpp -- the step of rods with the rods example's Cycle and
// NeighborSensing (bacterial_rods/src/main.rs:48-111), chili's order
// (run_sim.rs:259-289): forces (with the neighbour counts) -> mechanics update
// -> react_to_neighbors -> update_cycle -> update_cell_cycle_4 (division,
// update_cycle.rs:10-69) -> apply_boundary -> re-sort.
//
//   k_rod_mech          mechanics update in place (slot s -> slot s of the other
//                       buffer, no reflection), growth rate from the counts,
//                       spring-length growth, division flag
//   k_cycle_voxels      per voxel, in list order: identifiers of the mothers,
//                       counters of the daughters, daughters per voxel
//   (exclusive scan of the daughters per voxel)
//   k_expand_rod        re-layout into the grown order: every voxel's list
//                       followed by its daughters (the reference's
//                       cells.extend(new_cells)); RodMechanics::divide for mothers
//   k_grow_start        voxel starts of the grown order (division steps)
//   k_reflect_key_rod   per-vertex reflection, centroid voxel
//   (k_rank_ahead and the counting sort of the plain step, then k_advance_step)
// Each rod keeps the step of its first history push (`birth`): a daughter starts
// with an empty AuxStorage, its ring fills from the step after its division.
#pragma once

#include "kernels_rod.hpp"

namespace crm {

// The ring of a rod whose first push was at step `birth` (equal to ring_at when
// birth == 0): the entries are the pushes of steps max(birth, t + 1 - cap) .. t.
__device__ __forceinline__ Ring ring_rod(uint64_t t, uint64_t birth, uint32_t cap, uint32_t newest_first) {
    Ring r;
    r.cap = cap;
    r.wslot = cap ? (uint32_t)(t % cap) : 0u;
    const uint32_t size = (uint32_t)min<uint64_t>(t - birth + 1, cap);
    const uint32_t oldest = cap ? (uint32_t)((t + 1 - size) % cap) : 0u;
    r.n_read = (cap == 2 && size == 2) ? 2u : 1u;
#pragma unroll
    for (int i = 0; i < 2; i++)
        r.it[i] = cap == 0 ? 0u : (newest_first ? (oldest + size - 1 - i + 2 * cap) % cap : (oldest + i) % cap);
    return r;
}

struct CycleArgs {
    double base, threshold, saturation;
    uint32_t* divflag;  // [N]
    uint32_t* drank;    // [N]
    uint64_t* dctr;     // [N]
    uint32_t* ndiv;     // [Vt + 1]
    uint32_t* dbefore;  // [Vt + 1]
    uint64_t* vcount;   // [Vt]
    uint32_t* inst_div; // [n_inst] dividing rods per instance this step
    uint32_t n_inst;
};

__global__ void k_cycle_begin(Control* c, uint32_t* inst_div, uint32_t n_inst) {
    c->n_div = 0;
    for (uint32_t i = threadIdx.x; i < n_inst; i += blockDim.x) inst_div[i] = 0;
}

__global__ void k_advance_step(Control* c) { c->step += 1; }

// Voxel::id_counter after insert_initial_cells: the initial rods of each voxel
// (datastructures.rs:286-297).
__global__ __launch_bounds__(kBlock) void k_init_vcount(const uint32_t* start, uint64_t* vcount, uint32_t Vt) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g < Vt) vcount[g] = start[g + 1] - start[g];
}

// Mechanics update of lane (s, v) into the same slot of dst, then (first lane)
// react_to_neighbors + update_cycle (bacterial_rods/src/main.rs:66-73, 88-100).
template <int D>
__global__ __launch_bounds__(kBlock) void k_rod_mech(Arrays A, StateBuf src, StateBuf dst, RodArgs R, StepArgs sa,
                                                      CycleArgs Y) {
    __shared__ double win[kRodWinK];
    if (halted(A.ctrl, A.ctrl->step)) return;
    const BlockWork w = A.work[blockIdx.x];
    uint64_t s;
    uint32_t v;
    const bool active = rod_lane(w, R.nv, s, v);
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N, t = A.ctrl->step;
    const uint32_t nv = R.nv, K = nv * D;
    auto pos = [&](int i, int k) { return src.pos[(uint64_t)(i * D + k) * N + s]; };
    auto vel = [&](int i, int k) { return src.vel[(uint64_t)(i * D + k) * N + s]; };
    if (active && v + 2 < nv) {  // bending window v of the rod, shared with the next two lanes
        double F[D];
        rod_bend_window<D>(P, pos, (int)v, F);
#pragma unroll
        for (int k = 0; k < D; k++) win[threadIdx.x * 3 + k] = F[k];
    }
    __syncthreads();
    if (!active) return;
    const uint32_t id = src.id[s];
    const Ring rg = ring_rod(t, src.birth[s], sa.cap, sa.newest_first);
    auto wn = [&](int i, int k) { return win[(threadIdx.x - v + (uint32_t)i) * 3 + k]; };
    double x[D], vv[D], f[D], dx[D], dv[D], dxr[D], nx[D], nvel[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        x[k] = pos((int)v, k);
        vv[k] = vel((int)v, k);
        f[k] = A.force[(uint64_t)(v * D + k) * N + s];
    }
    rod_row_increment<D>(P, src.spring[s], nv, v, pos, vel, wn, f, dx, dv);
    rod_row_noise<D>(P, sa, id, t, nv, v, dxr);
    auto hist = [&](int i, int k, bool is_dv) {
        const double* h = is_dv ? src.hdv : src.hdx;
        return h[((uint64_t)rg.it[i] * K + v * D + k) * N + s];
    };
    rod_row_solve<D, true>(sa, rg, x, vv, dx, dv, dxr, hist, nx, nvel);
    for (uint32_t h = 0; h < rg.cap; h++) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            const uint64_t e = ((uint64_t)h * K + v * D + k) * N + s;
            dst.hdx[e] = h == rg.wslot ? dx[k] : src.hdx[e];
            dst.hdv[e] = h == rg.wslot ? dv[k] : src.hdv[e];
        }
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        dst.pos[(uint64_t)(v * D + k) * N + s] = nx[k];
        dst.vel[(uint64_t)(v * D + k) * N + s] = nvel[k];
    }
    if (v != 0) return;
    dst.species[s] = src.species[s];
    dst.id[s] = id;
    dst.vox[s] = src.vox[s];
    copy_rod_extras(src, s, dst, s, N, N);
    uint32_t n = 0;
    for (uint32_t q = 0; q < nv; q++) n += A.sense[s * nv + q];
    const double growth = n > 0 ? fmax((Y.base * (Y.saturation - (double)n)) / Y.saturation, 0.0) : Y.base;
    const double sl = src.spring[s] + growth * sa.dt;
    dst.growth[s] = growth;
    dst.spring[s] = sl;
    const bool divide = sl > Y.threshold;
    Y.divflag[s] = divide ? 1u : 0u;
    if (divide) {
        atomicAdd(&A.ctrl->n_div, 1u);
        atomicAdd(&Y.inst_div[w.inst], 1u);
    }
}

// update_cell_cycle_4's bookkeeping per voxel: mothers in list order get
// Division(voxel, ++counter) with their old identifier as parent; the daughters'
// counters follow all mothers'.
__global__ __launch_bounds__(kBlock) void k_cycle_voxels(Arrays A, StateBuf S, CycleArgs Y) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g > A.Vt) return;
    if (g == A.Vt) {
        Y.ndiv[g] = 0;  // the scan's last element: its exclusive sum is the total
        return;
    }
    // the voxel's plain index in its instance: the last instance whose first voxel
    // is at or below g (voxel offsets ascend with the instance)
    uint32_t lo = 0, hi = Y.n_inst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A.ip[mid].voxel_offset <= g) lo = mid;
        else hi = mid;
    }
    const uint64_t local = g - A.ip[lo].voxel_offset;
    const uint64_t N = A.N;
    const uint32_t b = A.start[g], e = A.start[g + 1];
    const uint64_t ctr = Y.vcount[g];
    uint32_t nd = 0;
    for (uint32_t s = b; s < e; s++) {
        if (!Y.divflag[s]) continue;
        Y.drank[s] = nd++;
#pragma unroll
        for (int j = 0; j < 3; j++) S.parent[j * N + s] = S.ident[j * N + s];
        S.ident[s] = 1;
        S.ident[N + s] = local;
        S.ident[2 * N + s] = ctr + nd;
    }
    for (uint32_t s = b; s < e; s++)
        if (Y.divflag[s]) Y.dctr[s] = ctr + nd + Y.drank[s] + 1;
    Y.vcount[g] = ctr + 2ull * nd;
    Y.ndiv[g] = nd;
}

// RodMechanics::divide (bacterial_rods.rs:586-662) of the rod in p (n vertices):
// m receives the mother's new vertices, d the daughter's; returns false on the
// reference's DivisionError.
template <int D>
__device__ bool rod_divide(const double* p, uint32_t n, double spring_length, double radius, double* m, double* d,
                           double& new_spring_length) {
    const double two = 1.0 + 1.0, one_half = 1.0 / two;
    const double div_factor = one_half - radius / ((double)n * spring_length);
    new_spring_length = div_factor * spring_length;
    double seg[CRM_MAX_ROD_VERTICES];
    for (uint32_t i = 0; i + 1 < n; i++) {
        double z[D];
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = p[i * D + k] - p[(i + 1) * D + k];
        seg[i] = sqrt(sum_squares_unfused<D>(z));
    }
    auto head = [&](uint32_t c) {
        double t = 0.0;
        for (uint32_t i = 0; i < c; i++) t = t + seg[i];
        return t;
    };
    auto tail = [&](uint32_t c) {
        double t = 0.0;
        for (uint32_t i = 0; i < c; i++) t = t + seg[n - 2 - i];
        return t;
    };
    const double seg_len = (head(n - 1) - two * radius) / (double)(n - 1) / two;
    for (uint32_t v = 0; v < n; v++) {
        const double L = (double)v * seg_len;
        int k = -1;
        for (uint32_t c = 0; c + 1 < n; c++)
            if (head(c) <= L) k = (int)c;
        if (k < 0) return false;
        const double q = (L - head((uint32_t)k)) / seg[k];
#pragma unroll
        for (int j = 0; j < D; j++) m[v * D + j] = p[k * D + j] * (1.0 - q) + p[(k + 1) * D + j] * q;
        int mm = -1;
        for (uint32_t c = 0; c < n; c++)
            if (tail(c) <= L) mm = (int)c;
        if (mm < 0 || (int)n - mm - 2 < 0) return false;
        const double pp = (L - tail((uint32_t)mm)) / seg[n - mm - 2];
#pragma unroll
        for (int j = 0; j < D; j++)
            d[(n - v - 1) * D + j] = p[(n - mm - 1) * D + j] * (1.0 - pp) + p[(n - mm - 2) * D + j] * pp;
    }
    return true;
}

// Re-layout into the grown order (thread per rod of the old order): rod s of
// voxel g moves to s + dbefore[g]; the daughters of g follow g's list at
// start[g + 1] + dbefore[g] + rank.  Strides: n_old in src, n_new in dst.
template <int D>
__global__ __launch_bounds__(kBlock) void k_expand_rod(Arrays A, StateBuf src, StateBuf dst, RodArgs R, StepArgs sa,
                                                        CycleArgs Y, uint64_t n_old, uint64_t n_new, double radius) {
    const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n_old) return;
    const uint32_t nv = R.nv, K = nv * D;
    const uint32_t g = src.vox[s];
    const uint64_t m = s + Y.dbefore[g];
    for (uint32_t c = 0; c < K; c++) {
        dst.pos[c * n_new + m] = src.pos[c * n_old + s];
        dst.vel[c * n_new + m] = src.vel[c * n_old + s];
        for (uint32_t h = 0; h < sa.cap; h++) {
            dst.hdx[(h * K + c) * n_new + m] = src.hdx[(h * K + c) * n_old + s];
            dst.hdv[(h * K + c) * n_new + m] = src.hdv[(h * K + c) * n_old + s];
        }
    }
    dst.species[m] = src.species[s];
    dst.id[m] = src.id[s];
    dst.vox[m] = g;
    copy_rod_extras(src, s, dst, m, n_old, n_new);
    if (!Y.divflag[s]) return;
    double p[CRM_MAX_ROD_VERTICES * D], pm[CRM_MAX_ROD_VERTICES * D], pd[CRM_MAX_ROD_VERTICES * D];
    for (uint32_t c = 0; c < K; c++) p[c] = src.pos[c * n_old + s];
    double nsl;
    const uint64_t t = A.ctrl->step;
    if (!rod_divide<D>(p, nv, src.spring[s], radius, pm, pd, nsl)) {
        raise_error(A.ctrl, CRM_DIVISION_ERROR, t, src.id[s]);
        return;
    }
    const uint64_t dslot = A.start[g + 1] + Y.dbefore[g] + Y.drank[s];
    for (uint32_t c = 0; c < K; c++) {
        dst.pos[c * n_new + m] = pm[c];
        dst.pos[c * n_new + dslot] = pd[c];
        dst.vel[c * n_new + dslot] = src.vel[c * n_old + s];
    }
    dst.spring[m] = nsl;
    dst.species[dslot] = src.species[s];
    dst.id[dslot] = (uint32_t)(n_old + Y.dbefore[g] + Y.drank[s]);
    dst.vox[dslot] = g;
    dst.spring[dslot] = nsl;
    dst.birth[dslot] = (uint32_t)(t + 1);  // a fresh AuxStorage: its first push is next step's
    dst.growth[dslot] = src.growth[s];
    dst.ident[dslot] = 1;
    dst.ident[n_new + dslot] = src.ident[n_old + s];  // the mother's voxel (set by k_cycle_voxels)
    dst.ident[2 * n_new + dslot] = Y.dctr[s];
#pragma unroll
    for (int j = 0; j < 3; j++) dst.parent[j * n_new + dslot] = src.parent[j * n_old + s];
}

// apply_boundary (per vertex, bacterial_rods.rs:489-533) in place, then the
// centroid voxel of every rod, as k_key_rod does for the plain step.
template <int D>
__global__ __launch_bounds__(kBlock) void k_reflect_key_rod(Arrays A, StateBuf S, RodArgs R) {
    __shared__ double newpos[kRodLdsK];
    __shared__ uint32_t outflag[kBlock];
    const BlockWork w = A.work[blockIdx.x];
    uint64_t s;
    uint32_t v;
    const bool active = rod_lane(w, R.nv, s, v);
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N, t = A.ctrl->step;
    const uint32_t nv = R.nv;
    if (active) {
        double nx[D], nvel[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            nx[k] = S.pos[(uint64_t)(v * D + k) * N + s];
            nvel[k] = S.vel[(uint64_t)(v * D + k) * N + s];
        }
        outflag[threadIdx.x] = rod_row_reflect<D, true>(P, nx, nvel) ? 1u : 0u;
#pragma unroll
        for (int k = 0; k < D; k++) {
            S.pos[(uint64_t)(v * D + k) * N + s] = nx[k];
            S.vel[(uint64_t)(v * D + k) * N + s] = nvel[k];
            newpos[threadIdx.x * 3 + k] = nx[k];
        }
    }
    __syncthreads();
    if (!active || v != 0) return;
    bool out = false;
    double c[D];
#pragma unroll
    for (int k = 0; k < D; k++) c[k] = 0.0;
    for (uint32_t q = 0; q < nv; q++) {
        out |= outflag[threadIdx.x + q] != 0;
#pragma unroll
        for (int k = 0; k < D; k++) c[k] = c[k] + newpos[(threadIdx.x + q) * 3 + k];
    }
    if (out) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, t, S.id[s]);
    uint32_t idx[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        ok = ok && axis_index(c[k] / (double)nv, P.min[ax], P.dx[ax], P.n[ax], idx[ax]);
    }
    uint32_t key;
    if (ok) {
        key = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2];
    } else {
        key = S.vox[s];
        atomicAdd((unsigned long long*)&A.ctrl->out_of_grid, 1ull);
    }
    A.key[s] = key;  // ranked by k_rank_ahead
}

// Voxel starts of the grown order: each voxel's list is followed by its daughters.
__global__ __launch_bounds__(kBlock) void k_grow_start(uint32_t* start, const uint32_t* dbefore, uint32_t Vt) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g <= Vt) start[g] += dbefore[g];
}

}  // namespace crm
