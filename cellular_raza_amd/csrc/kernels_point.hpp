// kernels_point.hpp -- gfx950 kernels of the per-step mechanics path for point agents.
//
// Step (run_sim.rs:380-403) as launched by the runtime:
//   k_force_point      pair loop over the 3^D voxel stencil (update_mechanics.rs:153-238)
//   k_integrate_point  Mechanics increment + Euler/AB step (solvers.rs:23-347), reflecting
//                      boundary (cartesian_cuboid_n.rs:594-631), next voxel index
//                      (cartesian_cuboid_n.rs:469-491) and its histogram
//   k_scan_*           exclusive scan of the voxel histogram            } integer counting sort,
//   k_scatter          slot -> sorted position                          } stable by previous slot
//   k_fixup            per-voxel order by previous slot, histogram reset } (sort_cells_in_voxels,
//   k_gather_point     permutes the state, writes force records          }  update_mechanics.rs:418-513)
//
// Agents live in SoA arrays sorted by voxel; a voxel's agents are the slots
// [start[v], start[v+1]).
#pragma once

#include "device_math.hpp"

namespace crm {

constexpr int kBlock = 256;

struct StateBuf {
    double* pos;  // [K][N]
    double* vel;  // [K][N]
    double* hdx;  // [2][K][N] Adams-Bashforth ring, position increments
    double* hdv;  // [2][K][N] velocity increments
    int32_t* species;
    uint32_t* id;   // agent index
    uint32_t* vox;  // global voxel of the slot's agent
};

struct Arrays {
    const InstParams* ip;
    const BlockWork* work;
    Control* ctrl;
    uint32_t* start;  // [Vt + 1]
    uint32_t* count;  // [Vt]
    uint32_t* key;    // [N] next voxel
    uint32_t* rank;   // [N] arrival rank inside the next voxel
    uint32_t* perm;   // [N] sorted slot -> previous slot
    double* force;    // [K][N]
    double4* rec64;   // [N] (x, y, z, species) absolute, FP64 path
    int4* recq;       // [N] (qx, qy, qz, species | outside << 31) fixed point in the voxel, mixed path
    uint64_t N;
    uint32_t Vt;
};

__device__ __forceinline__ void raise_error(Control* c, uint32_t kind, uint64_t step, uint64_t agent) {
    atomicOr(&c->error_flags, 1u << kind);
    unsigned long long k = ((unsigned long long)step << 32) | (unsigned long long)(agent & 0xffffffffu);
    atomicMin((unsigned long long*)&c->error_key, k);
}

template <typename T>
__device__ __forceinline__ const PotParams<T>& pot_of(const InstParams& P);
template <>
__device__ __forceinline__ const PotParams<double>& pot_of<double>(const InstParams& P) { return P.pd; }
template <>
__device__ __forceinline__ const PotParams<float>& pot_of<float>(const InstParams& P) { return P.pf; }

// Compensated (Kahan) running sum for the float pair path.
struct Kahan {
    float s, c;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// ---------------------------------------------------------------------------
// Force kernel.  One lane per agent; candidates are the agents of the clipped
// 3^D stencil (get_neighbor_voxel_indices, cartesian_cuboid_n.rs:641-662) plus
// the own voxel.  Weights, for the reference (one subdomain) mode:
//   same voxel:     1/2 f(i<-j by i) + 1/2 f(i<-j by j)   (update_mechanics.rs:186-203)
//   adjacent voxel: 1/2 f(i<-j by i)                        (update_mechanics.rs:90-106, 205-220)
// The physical mode gives adjacent pairs the same-voxel weight.  With a
// symmetric interaction both evaluations are equal and one is computed.
//
// Mixed precision: positions are 30-bit fixed point relative to the voxel
// origin (q = floor((p - origin) * 2^30 / dx)), so the difference of two agents
// in neighbouring voxels is an exact int32 and converts to float with one
// rounding: the direction is accurate to ~1e-7 however close the pair is.  A
// pair whose float distance lies in a narrow band around the cutoff radius (the
// forces jump there), a zero distance, or an agent flagged as outside its
// voxel's fixed-point range, is re-evaluated from the FP64 positions so that
// the cutoff decision is the reference's.
// ---------------------------------------------------------------------------

// blockIdx -> work block, so that each XCD (blocks are dealt round-robin over 8)
// walks one contiguous slab of agents and its L2 holds that slab's records.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t per = nb / 8u, rem = nb % 8u;
    const uint32_t x = b % 8u, k = b / 8u;
    return x < rem ? x * (per + 1u) + k : rem * (per + 1u) + (x - rem) * per + k;
}

constexpr int kQBits = 30;
constexpr int kQOne = 1 << kQBits;

template <int D, int POT, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_force_point(Arrays A, StateBuf S, uint32_t physical,
                                                          uint32_t symmetric) {
    const BlockWork w = A.work[xcd_block(blockIdx.x, gridDim.x)];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint32_t v = S.vox[s];
    const uint32_t lv = v - P.voxel_offset;
    uint32_t idx[3];
    idx[2] = lv % P.n[2];
    idx[1] = (lv / P.n[2]) % P.n[1];
    idx[0] = lv / (P.n[2] * P.n[1]);
    uint32_t lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = idx[a] == 0 ? 0 : idx[a] - 1;
        hi[a] = min(idx[a] + 2, P.n[a]);
    }
    using T = typename std::conditional<MIXED, float, double>::type;
    const PotParams<T>& PP = pot_of<T>(P);
    const T rc = cut_radius<POT, T>(PP);
    const float rc2 = (float)(rc * rc);
    const float band = 2e-6f * rc2;
    const uint64_t N = A.N;

    int qi[3] = {0, 0, 0};
    double me[3] = {0.0, 0.0, 0.0};
    int sp_i;
    uint32_t flag_i = 0;
    if constexpr (MIXED) {
        int4 r = A.recq[s];
        qi[0] = r.x;
        qi[1] = r.y;
        qi[2] = r.z;
        sp_i = r.w & 0x7fffffff;
        flag_i = (uint32_t)r.w >> 31;
    } else {
        double4 r = A.rec64[s];
        me[0] = r.x;
        me[1] = r.y;
        me[2] = r.z;
        sp_i = (int)r.w;
    }
    double accd[D];
    Kahan acck[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        accd[k] = 0.0;
        acck[k].s = 0.0f;
        acck[k].c = 0.0f;
    }
    bool calc_err = false;
    if (POT == CRM_POTENTIAL_NONE) hi[0] = lo[0];  // NoInteraction: no pair work
    for (uint32_t a = lo[0]; a < hi[0]; a++) {
        for (uint32_t b = lo[1]; b < hi[1]; b++) {
            const uint32_t row = P.voxel_offset + (a * P.n[1] + b) * P.n[2];
            uint32_t jb = A.start[row + lo[2]];
            for (uint32_t c = lo[2]; c < hi[2]; c++) {
                const uint32_t vv = row + c;
                const uint32_t je = A.start[vv + 1];
                if (jb != je) {
                    const bool full = physical || vv == v;
                    // fixed-point position of agent i relative to the candidate voxel's origin
                    int sh[D];
                    if constexpr (MIXED) {
                        const int off[3] = {(int)a - (int)idx[0], (int)b - (int)idx[1], (int)c - (int)idx[2]};
#pragma unroll
                        for (int k = 0; k < D; k++) sh[k] = qi[k] - off[k + 3 - D] * kQOne;
                    }
                    for (uint32_t j = jb; j < je; j++) {
                        if (j == s) continue;
                        T z[D];
                        int sp_j;
                        int cut = -1;
                        if constexpr (MIXED) {
                            int4 r = A.recq[j];
                            const int qj[3] = {r.x, r.y, r.z};
                            sp_j = r.w & 0x7fffffff;
                            float d2 = 0.0f;
#pragma unroll
                            for (int k = 0; k < D; k++) {
                                z[k] = (float)(sh[k] - qj[k]) * P.qunit[k + 3 - D];
                                d2 = d2 + z[k] * z[k];
                            }
                            const bool refine = (((uint32_t)r.w >> 31) | flag_i) || d2 == 0.0f ||
                                                fabsf(d2 - rc2) <= band;
                            if (refine) {
                                double zd[D], dd = 0.0;
#pragma unroll
                                for (int k = 0; k < D; k++) {
                                    zd[k] = S.pos[k * N + s] - S.pos[k * N + j];
                                    dd = dd + zd[k] * zd[k];
                                    z[k] = (float)zd[k];
                                }
                                cut = cut_decision<POT>(P.pd, sqrt(dd));
                            }
                        } else {
                            double4 r = A.rec64[j];
                            z[0] = me[0] - r.x;
                            z[1] = me[1] - r.y;
                            if constexpr (D == 3) z[2] = me[2] - r.z;
                            sp_j = (int)r.w;
                        }
                        T f[D];
                        calc_err |= point_force<POT, T, D>(z, sp_i, sp_j, PP, cut, f);
                        T wgt = T(0.5);
                        if (full) {
                            if (symmetric) {
                                wgt = T(1);
                            } else {
                                T zj[D], g[D];
#pragma unroll
                                for (int k = 0; k < D; k++) zj[k] = -z[k];
                                calc_err |= point_force<POT, T, D>(zj, sp_j, sp_i, PP, cut, g);
#pragma unroll
                                for (int k = 0; k < D; k++) {
                                    if constexpr (MIXED) acck[k].add(-g[k] * 0.5f);
                                    else accd[k] = accd[k] + (-g[k]) * 0.5;
                                }
                            }
                        }
#pragma unroll
                        for (int k = 0; k < D; k++) {
                            if constexpr (MIXED) acck[k].add(f[k] * wgt);
                            else accd[k] = accd[k] + f[k] * wgt;
                        }
                    }
                }
                jb = je;
            }
        }
    }
    if (calc_err) raise_error(A.ctrl, CRM_CALC_ERROR, A.ctrl->step, S.id[s]);
#pragma unroll
    for (int k = 0; k < D; k++) {
        double fk;
        if constexpr (MIXED) fk = (double)acck[k].s + (double)(-acck[k].c);
        else fk = accd[k];
        A.force[(uint64_t)k * N + s] = fk;
    }
}

// ---------------------------------------------------------------------------
// Integrator + boundary + next voxel.
// Ring semantics (chili/aux_storage.rs:41-130 with capacity = solver order):
// every step pushes (dx, dv) first, then reads the iterator (solvers.rs:178-186).
// t steps done before this one -> write slot t % cap; afterwards the ring holds
// min(t + 1, cap) entries, the oldest at slot (t + 1) % cap once full.
// ---------------------------------------------------------------------------
struct StepArgs {
    const double* noise;  // CRM_NOISE_BUFFER: this step's normals by agent index, else null
    uint64_t seed;
    double dt;
    double sqrt_dt;
    uint32_t noise_mode;
    uint32_t cap;           // solver order = ring capacity
    uint32_t newest_first;  // ring iteration order
    uint32_t pad;
};

template <int MECH, int D>
__device__ __forceinline__ void point_increment(const InstParams& P, const double* x, const double* v,
                                                const double* f, const double* z, bool have_noise, double dt,
                                                double sqrt_dt, double* dx, double* dv, double* dxr,
                                                double* dvr) {
#pragma unroll
    for (int k = 0; k < D; k++) dxr[k] = dvr[k] = 0.0;
    if constexpr (MECH == CRM_MECHANICS_NEWTON_DAMPED) {  // mechanics.rs:166-173
#pragma unroll
        for (int k = 0; k < D; k++) {
            dx[k] = v[k];
            dv[k] = f[k] / P.mass - P.damping * v[k];
        }
    } else if constexpr (MECH == CRM_MECHANICS_LANGEVIN) {  // mechanics.rs:434-463
#pragma unroll
        for (int k = 0; k < D; k++) {
            dx[k] = v[k];
            double dv1 = P.inv_mass * f[k];
            double dv2 = (-P.damping) * v[k];
            dv[k] = dv1 + dv2;
        }
        if (have_noise) {
#pragma unroll
            for (int k = 0; k < D; k++) dvr[k] = P.langevin_coeff * ((0.0 + sqrt_dt * z[k]) / dt);
        }
    } else {  // Brownian, mechanics.rs:336-357
#pragma unroll
        for (int k = 0; k < D; k++) {
            dx[k] = P.brownian_q * f[k];
            dv[k] = 0.0;
        }
        if (have_noise) {
#pragma unroll
            for (int k = 0; k < D; k++) dxr[k] = P.brownian_coeff * ((0.0 + sqrt_dt * z[k]) / dt);
        }
    }
}

template <int D, int MECH>
__global__ __launch_bounds__(kBlock) void k_integrate_point(Arrays A, StateBuf S, StepArgs sa) {
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    const uint64_t t = A.ctrl->step;
    const uint32_t id = S.id[s];

    double x[D], v[D], f[D], z[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        x[k] = S.pos[k * N + s];
        v[k] = S.vel[k * N + s];
        f[k] = A.force[k * N + s];
        z[k] = 0.0;
    }
    bool have_noise = false;
    if (sa.noise_mode == CRM_NOISE_BUFFER && sa.noise) {
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = sa.noise[(uint64_t)id * D + k];
        have_noise = true;
    } else if (sa.noise_mode == CRM_NOISE_PHILOX) {
        double g[4];
        gaussian_pair(sa.seed, id, t, 0, g[0], g[1]);
        if (D > 2) gaussian_pair(sa.seed, id, t, 1, g[2], g[3]);
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = g[k];
        have_noise = true;
    }
    double dx[D], dv[D], dxr[D], dvr[D];
    point_increment<MECH, D>(P, x, v, f, z, have_noise, sa.dt, sa.sqrt_dt, dx, dv, dxr, dvr);

    // ring push + iterator (see comment above)
    const uint32_t cap = sa.cap;
    double nx[D], nvel[D];
    if (cap == 0) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            nx[k] = (dx[k] * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            nvel[k] = (dv[k] * sa.dt + v[k]) * 1.0 + dvr[k] * sa.dt;
        }
    } else {
        const uint32_t wslot = (uint32_t)(t % cap);
        const uint32_t size = (uint32_t)min<uint64_t>(t + 1, cap);
        const uint32_t oldest = size < cap ? 0u : (uint32_t)((t + 1) % cap);
        uint32_t it[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
            it[i] = sa.newest_first ? (oldest + size - 1 - i + 2 * cap) % cap : (oldest + i) % cap;
        // which yielded element is the value pushed this step
        double h[2][D], hv[2][D];
        const uint32_t n_read = (cap == 2 && size == 2) ? 2u : 1u;
        for (uint32_t i = 0; i < n_read; i++) {
#pragma unroll
            for (int k = 0; k < D; k++) {
                if (it[i] == wslot) {
                    h[i][k] = dx[k];
                    hv[i][k] = dv[k];
                } else {
                    h[i][k] = S.hdx[((uint64_t)it[i] * D + k) * N + s];
                    hv[i][k] = S.hdv[((uint64_t)it[i] * D + k) * N + s];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            S.hdx[((uint64_t)wslot * D + k) * N + s] = dx[k];
            S.hdv[((uint64_t)wslot * D + k) * N + s] = dv[k];
        }
        if (n_read == 2) {  // adams_bashforth_3, solvers.rs:312-329
            const double f0 = 23.0 / 12.0, f1 = (-16.0) / 12.0, f2 = 5.0 / 12.0;
#pragma unroll
            for (int k = 0; k < D; k++) {
                double a = (dx[k] * f0 + (h[0][k] * f1 + h[1][k] * f2)) * sa.dt + x[k];
                nx[k] = a * 1.0 + dxr[k] * sa.dt;
                double b = (dv[k] * f0 + (hv[0][k] * f1 + hv[1][k] * f2)) * sa.dt + v[k];
                nvel[k] = b * 1.0 + dvr[k] * sa.dt;
            }
        } else {  // adams_bashforth_2, solvers.rs:331-347
            const double g0 = 3.0 / 2.0, g1 = (-1.0) / 2.0;
#pragma unroll
            for (int k = 0; k < D; k++) {
                nx[k] = ((dx[k] * g0 + h[0][k] * g1) * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
                nvel[k] = ((dv[k] * g0 + hv[0][k] * g1) * sa.dt + v[k]) * 1.0 + dvr[k] * sa.dt;
            }
        }
    }
    if (MECH == CRM_MECHANICS_BROWNIAN) {
#pragma unroll
        for (int k = 0; k < D; k++) nvel[k] = v[k];  // Brownian::set_velocity is a no-op
    }

    // apply_boundary (cartesian_cuboid_n.rs:594-631)
    bool out = false;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        if (nx[k] < P.min[ax]) {
            nx[k] = 2.0 * P.min[ax] - nx[k];
            nvel[k] = fabs(nvel[k]);
        }
        if (nx[k] > P.max[ax]) {
            nx[k] = 2.0 * P.max[ax] - nx[k];
            nvel[k] = -fabs(nvel[k]);
        }
        out |= (nx[k] < P.min[ax]) || (nx[k] > P.max[ax]);
    }
    if (out) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, t, id);

#pragma unroll
    for (int k = 0; k < D; k++) {
        S.pos[k * N + s] = nx[k];
        S.vel[k * N + s] = nvel[k];
    }

    // next voxel (get_index_of, cartesian_cuboid_n.rs:469-491); a point whose
    // index has no voxel keeps its voxel (update_mechanics.rs:430-453)
    uint32_t idx[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        ok = ok && axis_index(nx[k], P.min[ax], P.dx[ax], P.n[ax], idx[ax]);
    }
    uint32_t key;
    if (ok) {
        key = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2];
    } else {
        key = S.vox[s];
        atomicAdd((unsigned long long*)&A.ctrl->out_of_grid, 1ull);
    }
    A.key[s] = key;
    A.rank[s] = atomicAdd(&A.count[key], 1u);
}

// Initial binning (Domain::decompose, concepts-derive/src/domain.rs:357-380 with
// get_voxel_index_of_raw, cartesian_cuboid_n.rs:320-337): min < p strictly and
// the voxel must exist; agents keep input order inside a voxel.
template <int D>
__global__ __launch_bounds__(kBlock) void k_initial_key(Arrays A, StateBuf S, uint32_t n_vertices) {
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    double p[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        if (n_vertices == 1) {
            p[k] = S.pos[k * N + s];
        } else {  // rod centroid: row_sum / nrows (bacterial_rods.rs:368-372)
            double acc = 0.0;
            for (uint32_t q = 0; q < n_vertices; q++) acc = acc + S.pos[((uint64_t)q * D + k) * N + s];
            p[k] = acc / (double)n_vertices;
        }
    }
    uint32_t idx[3] = {0, 0, 0};
    bool strict = true, ok = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        strict = strict && (P.min[ax] < p[k]);
        double xx = (p[k] - P.min[ax]) / P.dx[ax];
        bool conv = (xx > -1.0) && (xx < 4294967296.0);
        double tt = conv ? trunc(xx) : 0.0;
        uint32_t i = tt <= 0.0 ? 0u : (uint32_t)tt;
        ok = ok && conv && i < P.n[ax];
        idx[ax] = conv ? min(i, P.n[ax] - 1) : 0u;
    }
    if (!strict) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, 0, S.id[s]);
    else if (!ok) raise_error(A.ctrl, CRM_INDEX_ERROR, 0, S.id[s]);
    const uint32_t key = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2];
    A.key[s] = key;
    A.rank[s] = atomicAdd(&A.count[key], 1u);
}

// ---------------------------------------------------------------------------
// Counting sort by voxel.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t& total) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) {
        uint32_t wt = wave_tot[q];
        if (q < wave) before += wt;
        all += wt;
    }
    __syncthreads();
    total = all;
    return before + inc - x;
}

__global__ __launch_bounds__(kBlock) void k_scan_local(const uint32_t* __restrict__ count,
                                                        uint32_t* __restrict__ start, uint32_t* __restrict__ bsum,
                                                        uint32_t Vt) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t q = base + i;
        vals[i] = q < Vt ? count[q] : 0u;
        sum += vals[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t q = base + i;
        if (q < Vt) start[q] = run;
        run += vals[i];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place; start[Vt] = total.
// advance != 0 also completes the step counter (every kernel that reads it for
// this step has run; k_gather_point reads the new value).
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t* __restrict__ bsum, uint32_t n_tiles,
                                                        uint32_t* __restrict__ start, uint32_t Vt, Control* ctrl,
                                                        uint32_t advance) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n_tiles; base += kBlock) {
        uint32_t q = base + threadIdx.x;
        uint32_t x = q < n_tiles ? bsum[q] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(x, total);
        if (q < n_tiles) bsum[q] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        start[Vt] = carry;
        if (advance) ctrl->step += 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* __restrict__ start, const uint32_t* __restrict__ bsum,
                                                      uint32_t Vt) {
    const uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < Vt) start[q] += bsum[q / kScanTile];
}

__global__ __launch_bounds__(kBlock) void k_scatter(Arrays A) {
    const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= A.N) return;
    A.perm[A.start[A.key[s]] + A.rank[s]] = (uint32_t)s;
}

// Orders each voxel's slots by previous slot (atomic arrival order is not
// deterministic) and clears the histogram for the next step.
__global__ __launch_bounds__(kBlock) void k_fixup(Arrays A) {
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= A.Vt) return;
    const uint32_t b = A.start[v], e = A.start[v + 1];
    A.count[v] = 0u;
    for (uint32_t i = b + 1; i < e; i++) {
        uint32_t x = A.perm[i];
        uint32_t j = i;
        while (j > b) {
            uint32_t y = A.perm[j - 1];
            if (y <= x) break;
            A.perm[j] = y;
            j--;
        }
        A.perm[j] = x;
    }
}

// Permutes the state into sorted order and writes the force records.  The ring
// holds min(step, cap) increments (k_scan_tiles has already advanced `step`).
template <int D, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_gather_point(Arrays A, StateBuf src, StateBuf dst, uint32_t cap) {
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    const uint32_t o = A.perm[s];
    const uint32_t n_hist = (uint32_t)min<uint64_t>(A.ctrl->step, cap);
    double p[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        p[k] = src.pos[k * N + o];
        dst.pos[k * N + s] = p[k];
        dst.vel[k * N + s] = src.vel[k * N + o];
    }
    for (uint32_t h = 0; h < n_hist; h++) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            dst.hdx[((uint64_t)h * D + k) * N + s] = src.hdx[((uint64_t)h * D + k) * N + o];
            dst.hdv[((uint64_t)h * D + k) * N + s] = src.hdv[((uint64_t)h * D + k) * N + o];
        }
    }
    const int32_t sp = src.species[o];
    dst.species[s] = sp;
    dst.id[s] = src.id[o];
    const uint32_t key = A.key[o];
    dst.vox[s] = key;
    if constexpr (MIXED) {
        const uint32_t lv = key - P.voxel_offset;
        uint32_t idx[3];
        idx[2] = lv % P.n[2];
        idx[1] = (lv / P.n[2]) % P.n[1];
        idx[0] = lv / (P.n[2] * P.n[1]);
        int q[3] = {0, 0, 0};
        uint32_t outside = 0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const int ax = k + 3 - D;
            const double origin = P.min[ax] + (double)idx[ax] * P.dx[ax];
            const double t = floor((p[k] - origin) * P.qinv[ax]);
            const bool ok = t >= 0.0 && t < (double)kQOne;  // false for NaN
            q[k] = ok ? (int)t : (t < 0.0 ? 0 : kQOne - 1);
            outside |= ok ? 0u : 1u;
        }
        A.recq[s] = make_int4(q[0], q[1], D == 3 ? q[2] : 0, (int)((uint32_t)sp | (outside << 31)));
    } else {
        A.rec64[s] = make_double4(p[0], p[1], D == 3 ? p[D - 1] : 0.0, (double)sp);
    }
}

}  // namespace crm
