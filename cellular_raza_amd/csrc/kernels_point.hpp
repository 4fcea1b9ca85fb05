// kernels_point.hpp -- gfx950 kernels of the per-step mechanics path for point agents.
//
// Step (run_sim.rs:380-403) as launched by the runtime:
//   k_force_tiled      pair loop over the 3^D voxel stencil (update_mechanics.rs:153-238)
//   k_key_point        next position (solvers.rs:23-347 + apply_boundary,
//                      cartesian_cuboid_n.rs:594-631) -> next voxel index
//                      (cartesian_cuboid_n.rs:469-491); NewtonDamped and
//                      Langevin get it from the previous gather instead
//   k_rank_ahead       ranks: stayers in order, arrivals atomic         } integer counting sort:
//   k_scan_local/tiles exclusive scan of the voxel histogram, its reset } stayers keep their order,
//   k_scatter          slot -> sorted position, final voxel starts      } movers appended by
//   k_gather_integrate_point  arrivals put in order of previous slot,   } previous slot (sort_cells_in_voxels,
//                      the full mechanics update written into the       }  update_mechanics.rs:418-513)
//                      agent's new sorted slot, and the force records
//   (set-up sorts and rods order the arrivals with k_fixup; systems whose voxels
//   fit one scan tile run k_sort_small, the scans and the scatter in one workgroup)
//
// Agents live in SoA arrays sorted by voxel; a voxel's agents are the slots
// [start[v], start[v+1]).
#pragma once

#include "device_math.hpp"

namespace crm {

constexpr int kBlock = 256;
// rank flag of an agent binned ahead that changes voxel (see k_rank_ahead)
constexpr uint32_t kMoved = 0x80000000u;

struct StateBuf {
    double* pos;  // [K][N]
    double* vel;  // [K][N]
    double* hdx;  // [2][K][N] Adams-Bashforth ring, position increments
    double* hdv;  // [2][K][N] velocity increments
    int32_t* species;
    uint32_t* id;   // agent index
    uint32_t* vox;  // global voxel of the slot's agent
    // rods only (null for points): RodMechanics::spring_length per rod, the step
    // of the rod's first history push (0, or the step after its division), and
    // the cycle state of the rods example
    double* spring;   // [N]
    uint32_t* birth;  // [N]
    double* growth;   // [N]
    uint64_t* ident;  // [3][N] CellIdentifier (kind, a, b)
    uint64_t* parent; // [3][N]
};

struct Arrays {
    const InstParams* ip;
    const BlockWork* work;
    Control* ctrl;
    uint32_t* start;  // [Vt + 1]
    uint32_t* start_local;  // [Vt] offset of each voxel inside its scan tile (k_scan_local)
    uint32_t* tile_off;     // [n_tiles] offset of each scan tile (k_scan_tiles)
    uint32_t* count;  // [Vt] arrivals per next voxel (atomic ranks)
    uint32_t* stay;   // [Vt] agents binned ahead that stay in their voxel (zero otherwise)
    uint32_t* key;    // [N] next voxel
    uint32_t* rank;   // [N] arrival rank inside the next voxel
    uint32_t* key_next;   // [N] next voxel of the following step (binned ahead by the gather)
    uint32_t* perm;   // [N] sorted slot -> previous slot
    double* force;    // [K][N]
    float* forcef;    // [K][N] in the same memory: point forces of the mixed path, stored in float
    double4* rec64;   // [N] (x, y, z, species) absolute, FP64 path
    int4* recq;       // [N] (qx, qy, qz, species | outside << 31) fixed point in the voxel, mixed path
    uint32_t* sense;  // [N] (rods: [N][nv]) neighbour counts by slot; null when sensing is off
    double sense_t;   // a pair is sensed when its squared distance is below this
    uint64_t N;
    uint32_t Vt;
};

// Squared norm summed in axis order with every product and sum rounded on its
// own (no fused multiply-add), as nalgebra's norm_squared computes it.
template <int D>
__device__ __forceinline__ double sum_squares_unfused(const double* z) {
#pragma clang fp contract(off)
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    return d2;
}

__device__ __forceinline__ void raise_error(Control* c, uint32_t kind, uint64_t step, uint64_t agent) {
    atomicOr(&c->error_flags, 1u << kind);
    unsigned long long k = ((unsigned long long)step << 32) | (unsigned long long)(agent & 0xffffffffu);
    atomicMin((unsigned long long*)&c->error_key, k);
}

// An earlier step than t raised an error.  The kernels of the steps after a
// failing one do nothing (the ones before the sort check this, the ones after it
// read Control::skip, set by k_scan_tiles), so the state stays at the end of the
// failing step, where run_simulation! returns the error (run_sim.rs:380-403).
__device__ __forceinline__ bool halted(const Control* c, uint64_t t) {
    return c->error_flags != 0 && (c->error_key >> 32) < t;
}

template <typename T>
__device__ __forceinline__ const PotParams<T>& pot_of(const InstParams& P);
template <>
__device__ __forceinline__ const PotParams<double>& pot_of<double>(const InstParams& P) { return P.pd; }
template <>
__device__ __forceinline__ const PotParams<float>& pot_of<float>(const InstParams& P) { return P.pf; }

// ---------------------------------------------------------------------------
// Force kernel.  One lane per agent; candidates are the agents of the clipped
// 3^D stencil (get_neighbor_voxel_indices, cartesian_cuboid_n.rs:641-662) plus
// the own voxel.  Weights, for the reference (one subdomain) mode:
//   same voxel:     1/2 f(i<-j by i) + 1/2 f(i<-j by j)   (update_mechanics.rs:186-203)
//   adjacent voxel: 1/2 f(i<-j by i)                        (update_mechanics.rs:90-106, 205-220)
// The physical mode gives adjacent pairs the same-voxel weight.  With a
// symmetric interaction both evaluations are equal and one is computed.
//
// Work unit: a tile of voxels -- plane a0, rows [b0, b1), columns [c0, c1) of
// one instance.  Its stencil region (planes a0-1..a0+1, rows b0-1..b1, columns
// c0-1..c1, clipped to the grid) is staged in LDS, one contiguous segment of
// sorted agents per (plane, row).
//
// Mixed precision: the force records hold 30-bit fixed-point positions relative
// to the agent's voxel origin (q = floor((p - origin) * 2^30 / dx)).  For the
// pair loop they are re-expressed relative to the tile region's origin with
// `bits` bits per voxel, bits = 31 - ceil(log2(region extent in voxels)), so that
// every coordinate and every difference is an exact int32: the displacement of
// two agents is one integer subtraction and one rounding to float.  A pair whose
// float distance lies in a narrow band around the cutoff radius (the forces jump
// there), a pair closer than 2e6 fixed-point steps (where the step would show in
// the direction), or an agent flagged as outside its voxel's fixed-point range,
// is re-evaluated from the FP64 positions, with the reference's cutoff decision.
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
// Record word w (mixed): species in bits 0-3, the voxel's column along the last
// axis in bits 4-30, the outside-its-voxel flag in bit 31.
constexpr uint32_t kSpeciesMask = 0xfu;
constexpr int kColShift = 4;
constexpr uint32_t kColMask = 0x7ffffffu;
constexpr uint32_t kFlagBit = 0x80000000u;

template <bool MIXED>
struct RecOf {
    using type = double4;
};
template <>
struct RecOf<true> {
    using type = int4;
};

__device__ __forceinline__ const double4* rec_array(const Arrays& A, std::false_type) { return A.rec64; }
__device__ __forceinline__ const int4* rec_array(const Arrays& A, std::true_type) { return A.recq; }

// A tile's stencil region: rows b0 - 1 .. b1 (RW of them, empty where outside the
// grid), columns [cl, ch) (NC of them) and, in 3-D, planes a0 - 1 .. a0 + 1 (NP = 3;
// NP = 1 in 2-D).
struct Region {
    uint32_t a0, b0;  // tile plane and first row
    uint32_t cl, ch;  // staged columns
    uint32_t RW, NC;  // staged rows (tile rows + 2), staged columns (ch - cl)
    int bits, shift;  // tile fixed point: bits per voxel, 30 - bits
};

// Voxel-relative record -> region-relative fixed point.  oa, ob: the record's
// plane and row offsets from the region origin (a0 - 1, b0 - 1).
template <int D>
__device__ __forceinline__ int4 to_region(const int4 r, int oa, int ob, const Region& G) {
    const int oc = (int)(((uint32_t)r.w >> kColShift) & kColMask) - (int)G.cl;
    const uint32_t w = (uint32_t)r.w & (kSpeciesMask | kFlagBit);
    if constexpr (D == 3)
        return make_int4((oa << G.bits) + (r.x >> G.shift), (ob << G.bits) + (r.y >> G.shift),
                         (oc << G.bits) + (r.z >> G.shift), (int)w);
    else
        return make_int4((ob << G.bits) + (r.x >> G.shift), (oc << G.bits) + (r.y >> G.shift), 0, (int)w);
}

struct Span {
    uint32_t b, e;
};

// Candidate order of an agent at (a, b, c): stencil rows b' ascending, then
// columns c' ascending, then planes a' ascending, then slot -- i.e. the 3^D voxels
// with the row outermost.  Both sources below produce this order.
//
// LDS source: the region is staged voxel by voxel in that same (row, column,
// plane) order, so the candidates of one stencil row (3 columns x all planes)
// are one contiguous LDS range and an agent's stream is three ranges.
struct Walk3 {
    uint32_t j, je;          // current range [., je)
    uint32_t j1, e1, j2, e2; // the ranges after it
    uint32_t left;           // how many of those remain
    // Moves past exhausted ranges, carrying the overshoot; false at the end.
    __device__ __forceinline__ bool settle() {
        while (j >= je) {
            if (left == 0) return false;
            const uint32_t over = j - je;
            j = j1 + over;
            je = e1;
            j1 = j2;
            e1 = e2;
            left--;
        }
        return true;
    }
};

template <int D, bool MIXED>
struct LdsSource {
    using Rec = typename RecOf<MIXED>::type;
    using W = Walk3;
    const Rec* lds;       // staged records, converted
    const uint16_t* vst;  // LDS index of the first agent of staged voxel v; vst[n_vox] = total
    Region G;
    uint32_t np, n_vox;
    const uint32_t* start;
    uint32_t voff, n1, n2;
    const uint16_t* sub;  // chili subdomain of each staged record's voxel (when the instance has several)
    __device__ __forceinline__ uint32_t sub_of(uint32_t j, const W&, const InstParams&) const { return sub[j]; }
    // staged voxel index of (row rb, column c, plane pa)
    __device__ __forceinline__ uint32_t vox(uint32_t rb, uint32_t c, uint32_t pa) const {
        return (rb * G.NC + (c - G.cl)) * np + pa;
    }
    __device__ __forceinline__ Rec record(uint32_t j, const W&) const { return lds[j]; }
    // global slot of LDS index j (only on the rare FP64 re-evaluation path)
    __device__ __forceinline__ uint32_t slot(uint32_t j, const W&) const {
        uint32_t lo = 0, hi = n_vox;  // largest v with vst[v] <= j
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint32_t)vst[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t pa = lo % np, t = lo / np, c = G.cl + t % G.NC, rb = t / G.NC;
        const uint32_t a = G.a0 + pa - (np == 3 ? 1u : 0u), b = G.b0 - 1 + rb;
        return start[voff + (a * n1 + b) * n2 + c] + (j - (uint32_t)vst[lo]);
    }
};

// Global source (tiles whose region does not fit in LDS): the stream walks the
// stencil voxel by voxel in the same order.
struct WalkVox {
    uint32_t j, je;
    uint32_t rb, cc, pa;     // current voxel: row b_lo + rb, column c_lo + cc, plane a_lo + pa
    uint32_t nb, ncol, na;   // stencil extent (clipped)
    uint32_t a_lo, b_lo, c_lo;
    const uint32_t* start;
    uint32_t voff, n1, n2;
    __device__ __forceinline__ uint32_t gvox() const {
        return voff + ((a_lo + pa) * n1 + (b_lo + rb)) * n2 + c_lo + cc;
    }
    __device__ __forceinline__ bool settle() {
        while (j >= je) {
            const uint32_t over = j - je;
            if (++pa == na) {
                pa = 0;
                if (++cc == ncol) {
                    cc = 0;
                    if (++rb == nb) return false;
                }
            }
            const uint32_t g = gvox();
            j = start[g] + over;
            je = start[g + 1];
        }
        return true;
    }
};

template <int D, bool MIXED>
struct GlobalSource {
    using Rec = typename RecOf<MIXED>::type;
    using W = WalkVox;
    const Rec* rec;
    Region G;
    __device__ __forceinline__ Rec record(uint32_t j, const W& w) const {
        if constexpr (MIXED)
            return to_region<D>(rec[j], (int)(w.a_lo + w.pa + 1 - G.a0), (int)(w.b_lo + w.rb + 1 - G.b0), G);
        else
            return rec[j];
    }
    __device__ __forceinline__ uint32_t slot(uint32_t j, const W&) const { return j; }
    __device__ __forceinline__ uint32_t sub_of(uint32_t, const W& w, const InstParams& P) const {
        return subdomain_of(P, w.gvox() - w.voff);
    }
};

// Strength row of the species matrix in four scalars (a register array indexed
// by a run-time species would be placed in scratch memory).
template <typename T>
struct Row4 {
    T v0, v1, v2, v3;
    // bit tests rather than a compare chain on k, which the compiler turns back into
    // an indexed (scratch) table
    __device__ __forceinline__ T pick(int k) const {
        const bool b0 = (k & 1) != 0, b1 = (k & 2) != 0;
        const T lo = b0 ? v1 : v0, hi = b0 ? v3 : v2;
        return b1 ? hi : lo;
    }
};

// Force on the agent in slot s: `me` is its record in the source's coordinates,
// `self` its index in the source's index space, `ov` its own voxel's range there
// (same-voxel pairs) and `w0` the start of its candidate stream.
//
// Two passes per chunk of 64 stream positions: a cheap squared-distance test
// marks the candidates that can be inside the cutoff in a 64-bit mask, then the
// chunk is replayed and only the marked pairs are evaluated, in stream order.
// Every unmarked pair is beyond the cutoff, where each potential's force is
// exactly zero, so the sums are those of the full loop without its zero terms.
// The wave pays max-over-lanes of the marked count for the expensive pass
// instead of max-over-lanes of the candidate count.
//
// Lanes per agent (LDS source): 2^lg consecutive lanes of a wave share one agent;
// lane `sub` of the group walks the sub-th of 2^lg equal parts of the stream, and
// the group adds its partial sums with wavefront shuffles (a butterfly, so every
// lane of the group holds the same total) before lane 0 writes the force.
template <int D, int POT, bool MIXED, bool SENSE, typename Src>
__device__ __forceinline__ void agent_force(const Arrays& A, const StateBuf& S, const InstParams& P, uint64_t s,
                                            uint32_t self, const typename RecOf<MIXED>::type me, Span ov,
                                            typename Src::W cur, uint32_t physical, uint32_t symmetric,
                                            const Src& src, const float* unit, uint32_t sub = 0, uint32_t lg = 0) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using Rec = typename RecOf<MIXED>::type;
    using W = typename Src::W;
    const uint64_t N = A.N;
    const PotParams<T>& PP = pot_of<T>(P);
    const T rc = cut_radius<POT, T>(PP);
    const float rc2 = (float)(rc * rc);
    const float band = 2e-6f * rc2;

    int sp_i;
    bool flag_i = false;
    if constexpr (MIXED) {
        sp_i = (int)((uint32_t)me.w & kSpeciesMask);
        flag_i = me.w < 0;
    } else {
        sp_i = (int)me.w;
    }
    // candidate test thresholds: mixed marks d2 <= rc2 + band (the band pairs are
    // re-decided in FP64) and every pair with a flagged agent; FP64 marks
    // d2 <= rc^2 (1 + 1e-12), beyond which the rounded sqrt(d2) exceeds rc.
    // Mixed pass 1 measures in units of the first axis' fixed-point step (the
    // other axes scaled by their step ratio) and keeps half a band more margin.
    float r1 = 1.0f, r2 = 1.0f, thr_f = 0.0f, close2 = 0.0f;
    if constexpr (MIXED) {
        const float u0 = unit[0];
        r1 = unit[1] / u0;
        r2 = unit[D - 1] / u0;
        thr_f = (rc2 + 1.5f * band) / (u0 * u0);
        if (!(thr_f >= 0.0f) || flag_i) thr_f = __builtin_inff();
        // pairs closer than 2e6 fixed-point steps (where the quantised direction
        // would be off by more than 1e-6) take their displacement from FP64
        const float umax = fmaxf(fmaxf(unit[0], unit[1]), unit[D - 1]);
        close2 = (2e6f * umax) * (2e6f * umax);
    }
    const double rcd = (double)cut_radius<POT, double>(P.pd);
    double thr_d = rcd * rcd * (1.0 + 1e-12);
    if (!(thr_d >= 0.0)) thr_d = __builtin_inf();

    // Neighbour sensing (NeighborSensing::accumulate_information, concepts
    // interaction.rs:64-79, called per stencil pair at update_mechanics.rs:105,
    // 199-202) counts the candidates with squared distance below sense_t, the
    // smallest double whose square root reaches the radius, so that the test is
    // the reference's norm < radius.  Pass 1 then marks up to the larger of the
    // two radii, and pass 2 replays pass 1's force test for each marked pair.
    // Mixed precision counts from the fixed-point distance outside a band that
    // bounds its error (quantisation of both ends plus float rounding) and
    // decides the band's pairs from the FP64 positions.
    constexpr bool sense = SENSE;
    uint32_t n_sense = 0;
    float mark_f = thr_f, rs2 = 0.0f, band_s = 0.0f;
    double mark_d = thr_d;
    if constexpr (SENSE) {
        if constexpr (MIXED) {
            const float umax = fmaxf(fmaxf(unit[0], unit[1]), unit[D - 1]);
            const float q = 2.0f * (float)D * umax;
            rs2 = (float)A.sense_t;
            band_s = 2e-6f * rs2 + 4.0f * sqrtf(rs2) * q + q * q;
            float t = (rs2 + 2.0f * band_s) / (unit[0] * unit[0]);
            if (!(t >= 0.0f) || flag_i) t = __builtin_inff();
            mark_f = fmaxf(thr_f, t);
        } else {
            mark_d = fmax(thr_d, A.sense_t * (1.0 + 1e-12));
        }
    }
    const bool wider = MIXED ? mark_f > thr_f : mark_d > thr_d;
    // chili with several subdomains: pairs across a subdomain border get both halves
    // (update_mechanics.rs:285-382)
    const bool subs = P.sub_on != 0;
    const uint32_t my_sub = subs ? subdomain_of(P, S.vox[s] - P.voxel_offset) : 0u;

    // species-matrix row (own species) and column (own species as ext): Morse
    // strengths, or the Langevin sorting interaction's attractions
    Row4<T> srow, scol;
    {
        T r4[CRM_MAX_SPECIES], c4[CRM_MAX_SPECIES];
#pragma unroll
        for (int k = 0; k < CRM_MAX_SPECIES; k++) {
            if constexpr (POT == CRM_POTENTIAL_MORSE_SPECIES) {
                r4[k] = PP.strength * PP.species_matrix[sp_i * CRM_MAX_SPECIES + k];
                c4[k] = PP.strength * PP.species_matrix[k * CRM_MAX_SPECIES + sp_i];
            } else if constexpr (POT == CRM_POTENTIAL_CELL_SORTING_LANGEVIN) {  // attraction
                r4[k] = PP.species_matrix[sp_i * CRM_MAX_SPECIES + k];
                c4[k] = PP.species_matrix[k * CRM_MAX_SPECIES + sp_i];
            } else {
                r4[k] = c4[k] = PP.strength;
            }
        }
        srow = Row4<T>{r4[0], r4[1], r4[2], r4[3]};
        scol = Row4<T>{c4[0], c4[1], c4[2], c4[3]};
    }
    // per-agent sums: double in FP64 mode; float in mixed mode, whose forces are
    // stored as float (a compensated Kahan sum changed the stored forces by
    // nothing measurable -- 1M-agent 3-D sorting error 4.29e-7 either way, 4M-agent
    // Mie lattice 1.91e-6 against 1.93e-6 -- and cost 5 % of the kernel; a double
    // accumulator was slower)
    double accd[D];
    float ks[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        accd[k] = 0.0;
        ks[k] = 0.0f;
    }
    auto kadd = [&](int k, float x) { ks[k] = ks[k] + x; };
    bool calc_err = false;

    // Pair displacement own - candidate; returns the squared distance.
    auto displacement = [&](const Rec& rj, T* z) -> T {
        T d2 = T(0);
        if constexpr (MIXED) {
            z[0] = (float)(me.x - rj.x) * unit[0];
            z[1] = (float)(me.y - rj.y) * unit[1];
            if constexpr (D == 3) z[2] = (float)(me.z - rj.z) * unit[2];
        } else {
            z[0] = me.x - rj.x;
            z[1] = me.y - rj.y;
            if constexpr (D == 3) z[2] = me.z - rj.z;
        }
#pragma unroll
        for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
        return d2;
    };

    // Pass 1's squared distance: mixed in units of the first axis' fixed-point
    // step, FP64 with fused products.
    auto mark_dist = [&](const Rec& rj) -> T {
        if constexpr (MIXED) {
            const float c0 = (float)(me.x - rj.x);
            const float c1 = (float)(me.y - rj.y) * r1;
            float d2 = fmaf(c1, c1, c0 * c0);
            if constexpr (D == 3) {
                const float c2 = (float)(me.z - rj.z) * r2;
                d2 = fmaf(c2, c2, d2);
            }
            return d2;
        } else {
            const double z0 = me.x - rj.x, z1 = me.y - rj.y;
            double d2 = fma(z1, z1, z0 * z0);
            if constexpr (D == 3) {
                const double z2 = me.z - rj.z;
                d2 = fma(z2, z2, d2);
            }
            return d2;
        }
    };

    // The pair with the candidate at walk position w, weighted as in the reference.
    auto evaluate = [&](const W& w) {
        const uint32_t j = w.j;
        const Rec rj = src.record(j, w);
        T z[D];
        const T d2 = displacement(rj, z);
        if constexpr (SENSE) {
            bool in;
            if constexpr (MIXED) {
                if (flag_i || rj.w < 0 || fabsf(d2 - rs2) <= band_s) {
                    const uint32_t gj = src.slot(j, w);
                    double zd[D];
#pragma unroll
                    for (int k = 0; k < D; k++) zd[k] = S.pos[k * N + s] - S.pos[k * N + gj];
                    in = sum_squares_unfused<D>(zd) < A.sense_t;
                } else {
                    in = d2 < rs2;
                }
            } else {
                in = sum_squares_unfused<D>(z) < A.sense_t;
            }
            n_sense += in ? 1u : 0u;
            // the force only for the pairs pass 1 marks without sensing (all of
            // them unless the sensing radius widened the marks)
            if constexpr (POT == CRM_POTENTIAL_NONE) return;
            if (wider) {
                bool fmark;
                if constexpr (MIXED) fmark = mark_dist(rj) <= thr_f || rj.w < 0;
                else fmark = mark_dist(rj) <= thr_d;
                if (!fmark) return;
            }
        }
        const bool full = physical || (j >= ov.b && j < ov.e) || (subs && src.sub_of(j, w, P) != my_sub);
        int sp_j;
        int cut = -1;
        if constexpr (MIXED) {
            sp_j = (int)((uint32_t)rj.w & kSpeciesMask);
            const bool refine = flag_i || rj.w < 0 || d2 < close2 || fabsf(d2 - rc2) <= band;
            if (refine) {
                const uint32_t gj = src.slot(j, w);
                double zd[D], dd = 0.0;
#pragma unroll
                for (int k = 0; k < D; k++) {
                    zd[k] = S.pos[k * N + s] - S.pos[k * N + gj];
                    dd = dd + zd[k] * zd[k];
                    z[k] = (float)zd[k];
                }
                cut = cut_decision<POT>(P.pd, sqrt(dd));
            }
        } else {
            sp_j = (int)rj.w;
        }
        T f[D];
        const bool same_sp = sp_i == sp_j;
        calc_err |= point_force<POT, T, D>(z, same_sp, srow.pick(sp_j), PP, cut, f);
        T wgt = T(0.5);
        if (full) {
            if (symmetric) {
                wgt = T(1);
            } else {
                T zj[D], g[D];
#pragma unroll
                for (int k = 0; k < D; k++) zj[k] = -z[k];
                calc_err |= point_force<POT, T, D>(zj, same_sp, scol.pick(sp_j), PP, cut, g);
#pragma unroll
                for (int k = 0; k < D; k++) {
                    if constexpr (MIXED) kadd(k, -g[k] * 0.5f);
                    else accd[k] = accd[k] + (-g[k]) * 0.5;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            if constexpr (MIXED) kadd(k, f[k] * wgt);
            else accd[k] = accd[k] + f[k] * wgt;
        }
    };

    if constexpr (std::is_same<W, Walk3>::value) {
        // LDS source: the stream is three ranges of LDS indices, [j0, e0), [j1, e1),
        // [j2, e2), read as one index space p in [0, L): p < B1 is range 0, p < B2
        // range 1, the rest range 2 (empty ranges fold away), so the walk needs no
        // per-candidate range bookkeeping.  The own record sits in range 1 at p_self
        // and is unmarked once per chunk.  Pass 1 marks, pass 2 evaluates the marked
        // positions in stream order -- the same pairs, order and arithmetic as the
        // generic walk below.
        const bool active = POT != CRM_POTENTIAL_NONE || sense;
        const uint32_t B1 = cur.je - cur.j, B2 = B1 + (cur.e1 - cur.j1);
        const uint32_t L = active ? B2 + (cur.e2 - cur.j2) : 0u;
        const uint32_t o1 = cur.j1 - cur.je, o2 = cur.j2 - cur.e1;  // modular offsets
        // this lane's part [pb, pe) of the stream, walked by the offset r = p - pb
        // (the same for every lane of a wave, so the loop counters stay scalar)
        const uint32_t part = (L + (1u << lg) - 1u) >> lg;
        const uint32_t pb = min(L, sub * part), Lr = min(L, pb + part) - pb;
        const uint32_t j0 = cur.j + pb + (pb >= B1 ? o1 : 0u) + (pb >= B2 ? o2 : 0u);
        // range boundaries ahead of the part's start, and their offsets
        const uint32_t R1 = B1 > pb ? B1 - pb : 0u, R2 = B2 > pb ? B2 - pb : 0u;
        const uint32_t off1 = pb < B1 ? o1 : 0u, off2 = pb < B2 ? o2 : 0u;
        const uint32_t r_self = B1 + (self - cur.j1) - pb;  // wraps above Lr when before the part
        // LDS byte offset of stream offset r (the range offsets scaled once, not per
        // candidate), its record and its LDS index
        constexpr uint32_t RB = (uint32_t)sizeof(Rec);
        const uint32_t jb0 = j0 * RB, off1b = off1 * RB, off2b = off2 * RB;
        const char* lds_b = reinterpret_cast<const char*>(src.lds);
        auto atb = [&](uint32_t r) -> uint32_t { return jb0 + r * RB + (r >= R1 ? off1b : 0u) + (r >= R2 ? off2b : 0u); };
        auto rec_at = [&](uint32_t r) -> Rec { return *reinterpret_cast<const Rec*>(lds_b + atb(r)); };
        auto at = [&](uint32_t r) -> uint32_t { return atb(r) / RB; };
        // unit grids (equal fixed-point steps on every axis) need no axis scaling in
        // pass 1 (r1 = r2 = 1 exactly, so the marks are the same)
        const bool cubic = !MIXED || (r1 == 1.0f && r2 == 1.0f);
        auto marked = [&](auto cub, uint32_t p) -> uint32_t {
            const Rec rj = rec_at(p);
            bool mark;
            if constexpr (MIXED) {
                float d2;
                if constexpr (decltype(cub)::value) {
                    const float c0 = (float)(me.x - rj.x), c1 = (float)(me.y - rj.y);
                    d2 = fmaf(c1, c1, c0 * c0);
                    if constexpr (D == 3) {
                        const float c2 = (float)(me.z - rj.z);
                        d2 = fmaf(c2, c2, d2);
                    }
                } else {
                    d2 = mark_dist(rj);
                }
                mark = d2 <= mark_f || rj.w < 0;
            } else {
                mark = mark_dist(rj) <= mark_d;
            }
            return mark ? 1u : 0u;
        };
        auto pass1 = [&](auto cub, uint32_t p0, uint32_t qn, uint32_t& mlo, uint32_t& mhi) {
            const uint32_t q1 = min(qn, 32u);
            for (uint32_t q = 0; q < q1; q++) mlo |= marked(cub, p0 + q) << q;
            for (uint32_t q = 32; q < qn; q++) mhi |= marked(cub, p0 + q) << (q - 32u);
        };
        for (uint32_t p0 = 0; p0 < Lr; p0 += 64u) {
            const uint32_t qn = min(64u, Lr - p0);
            uint32_t mlo = 0, mhi = 0;
            if (cubic) pass1(std::true_type{}, p0, qn, mlo, mhi);
            else pass1(std::false_type{}, p0, qn, mlo, mhi);
            const uint32_t ds = r_self - p0;  // the self pair is never evaluated
            if (ds < 32u) mlo &= ~(1u << ds);
            else if (ds < 64u) mhi &= ~(1u << (ds - 32u));
            W w = cur;
            while (mlo | mhi) {
                uint32_t q;
                if (mlo) {
                    q = (uint32_t)__builtin_ctz(mlo);
                    mlo &= mlo - 1;
                } else {
                    q = 32u + (uint32_t)__builtin_ctz(mhi);
                    mhi &= mhi - 1;
                }
                w.j = at(p0 + q);
                evaluate(w);
            }
        }
    } else {
    bool live = POT != CRM_POTENTIAL_NONE || sense;
    if (live) live = cur.settle();
    while (live) {
        const W c0 = cur;
        // pass 1: mark the candidates that can be inside the cutoff (bit p of
        // mlo / mhi for stream position p of the chunk)
        uint32_t mlo = 0, mhi = 0;
        for (uint32_t p = 0; p < 64u; p++) {
            const uint32_t j = cur.j;
            const Rec rj = src.record(j, cur);
            bool mark;
            if constexpr (MIXED) mark = mark_dist(rj) <= mark_f || rj.w < 0;
            else mark = mark_dist(rj) <= mark_d;
            mark = mark && j != self;
            const uint32_t bit = 1u << (p & 31u);
            if (p < 32u) mlo = mark ? (mlo | bit) : mlo;
            else mhi = mark ? (mhi | bit) : mhi;
            ++cur.j;
            live = cur.settle();
            if (!live) break;
        }
        // pass 2: replay the chunk and evaluate the marked pairs in stream order
        W w = c0;
        uint32_t pos = 0;
        while (mlo | mhi) {
            uint32_t p;
            if (mlo) {
                p = (uint32_t)__builtin_ctz(mlo);
                mlo &= mlo - 1;
            } else {
                p = 32u + (uint32_t)__builtin_ctz(mhi);
                mhi &= mhi - 1;
            }
            w.j += p - pos;
            pos = p;
            w.settle();
            evaluate(w);
        }
    }
    }
    double fsum[D];
#pragma unroll
    for (int k = 0; k < D; k++) fsum[k] = MIXED ? (double)ks[k] : accd[k];
    if (lg) {  // the group's partial sums, by xor butterfly (same order on every lane)
        uint32_t err = calc_err ? 1u : 0u;
        for (uint32_t o = 1; o < (1u << lg); o <<= 1) {
#pragma unroll
            for (int k = 0; k < D; k++) fsum[k] = fsum[k] + __shfl_xor(fsum[k], (int)o, 64);
            err |= (uint32_t)__shfl_xor((int)err, (int)o, 64);
            if constexpr (SENSE) n_sense += (uint32_t)__shfl_xor((int)n_sense, (int)o, 64);
        }
        calc_err = err != 0;
        if (sub) return;
    }
    if (calc_err) raise_error(A.ctrl, CRM_CALC_ERROR, A.ctrl->step, S.id[s]);
    // the domain force follows the pair forces (calculate_custom_domain_force after
    // the subdomain's pairs, run_sim.rs:277-279, update_mechanics.rs:244-274)
    const bool spring = P.dforce == CRM_DOMAIN_FORCE_SPRING;
#pragma unroll
    for (int k = 0; k < D; k++) {
        double f = fsum[k];
        if (spring) f = f + (-P.dk[k]) * (S.pos[(uint64_t)k * N + s] - P.dc[k]);
        if constexpr (MIXED) A.forcef[(uint64_t)k * N + s] = (float)f;  // the pair sums are float already
        else A.force[(uint64_t)k * N + s] = f;
    }
    if constexpr (SENSE) A.sense[s] = n_sense;
}

// A force tile: plane a0, rows [b0, b1), columns [c0, c1) of one instance.
constexpr int kTileRowsMax = 29;
constexpr int kLdsRecordsMixed = 1024;  // 16 KiB of int4
constexpr int kLdsRecordsFP64 = 896;    // 28 KiB of double4
constexpr int kLdsVoxelStarts = 768;    // 1.5 KiB of voxel-start indices (staged voxels + 1)
constexpr int kOwnVoxMax = 256;         // own voxels per tile
constexpr int kSortBins = 64;           // candidate-count classes for the lane order

struct ForceTile {
    uint32_t inst, a0, b0, b1, c0, c1;
};

// An own agent of a force tile, looked up once: its record's LDS index, its
// voxel's LDS range, the three LDS ranges of its candidate stream, its slot.
struct OwnEntry {
    uint16_t self, ovb, ove;
    uint16_t r[6];
    uint16_t pad;
    uint32_t slot;
};

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += y;
    }
    return v;
}

// Exclusive prefix sum over the workgroup; total = the sum of all x.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t& total) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    const int wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(x);
    if ((threadIdx.x & 63) == 63) wave_tot[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) {
        const uint32_t wt = wave_tot[q];
        if (q < wave) before += wt;
        all += wt;
    }
    __syncthreads();
    total = all;
    return before + inc - x;
}

// a / d for a < 2^14, 0 < d <= 64, with d's float reciprocal: (a + 1/2) / d keeps a
// fractional part >= 1/128 that the two float roundings cannot carry across.
__device__ __forceinline__ uint32_t small_div(uint32_t a, float inv_d) {
    return (uint32_t)(((float)a + 0.5f) * inv_d);
}

// Workgroup per tile: stage the region voxel by voxel (converted records, voxel
// starts) in LDS, then one lane per own agent sums its forces from LDS.  A tile
// whose region does not fit (local crowding) reads global memory instead; the
// candidate order and the arithmetic are the same.
// MULTI: the tiling policy gives each agent 2^lanes_log2 lanes (a separate
// instantiation, so that the one-lane kernel carries none of the group logic).
template <int D, int POT, bool MIXED, bool SENSE, bool MULTI>
__global__ __launch_bounds__(kBlock, MIXED ? 5 : 4) void k_force_tiled(Arrays A, StateBuf S,
                                                                     const ForceTile* __restrict__ tiles,
                                                                     uint32_t physical, uint32_t symmetric,
                                                                     uint32_t no_lds, uint32_t rank,
                                                                     uint32_t lanes_log2) {
    const uint32_t lg = MULTI ? lanes_log2 : 0u;
    using Rec = typename RecOf<MIXED>::type;
    constexpr int CAP = MIXED ? kLdsRecordsMixed : kLdsRecordsFP64;
    constexpr uint32_t NP = D == 3 ? 3u : 1u;         // staged planes
    const ForceTile T = tiles[xcd_block(blockIdx.x, gridDim.x)];  // loaded beside the control block
    if (halted(A.ctrl, A.ctrl->step)) return;
    constexpr uint32_t PA_OWN = D == 3 ? 1u : 0u;     // the tile's own plane among them
    __shared__ Rec lds[CAP];
    __shared__ uint16_t vst[kLdsVoxelStarts];        // LDS index of each staged voxel's first agent
    __shared__ uint16_t own_pre[kOwnVoxMax + 1];     // own agents before own voxel u
    __shared__ uint32_t own_gs[kOwnVoxMax];          // global slot of own voxel u's first agent
    __shared__ uint8_t ord[kBlock];                  // own agents in lane order
    __shared__ uint32_t bins[kSortBins];
    __shared__ OwnEntry own_tab[kBlock];             // own agent lookups (sorted case)
    __shared__ uint16_t lsub[CAP];                   // chili subdomain of each staged record (sub_on only)
    const Rec* __restrict__ rec = rec_array(A, std::integral_constant<bool, MIXED>());
    const InstParams& P = A.ip[T.inst];
    const uint32_t n0 = P.n[0], n1 = P.n[1], n2 = P.n[2], voff = P.voxel_offset;
    Region G;
    G.a0 = T.a0;
    G.b0 = T.b0;
    G.RW = T.b1 - T.b0 + 2;
    G.cl = T.c0 == 0 ? 0u : T.c0 - 1u;
    G.ch = min(T.c1 + 1u, n2);
    G.NC = G.ch - G.cl;
    {
        const uint32_t ext = max(max(NP, G.RW), G.NC);
        const int lg = ext <= 1 ? 0 : 32 - __clz(ext - 1);  // ceil(log2(ext))
        G.bits = min(30, 31 - lg);
        G.shift = 30 - G.bits;
    }
    float unit[3];
#pragma unroll
    for (int k = 0; k < D; k++) unit[k] = MIXED ? ldexpf(P.qunit[k + 3 - D], G.shift) : 0.0f;
    const uint32_t R = T.b1 - T.b0, W = T.c1 - T.c0;
    const uint32_t n_vox = G.RW * G.NC * NP;
    const uint32_t n_ownv = R * W;
    const bool fits = !no_lds && n_vox < (uint32_t)kLdsVoxelStarts && n_ownv <= (uint32_t)kOwnVoxMax;

    // The step's re-binning ranks for the tile's own voxels (k_rank_ahead's work;
    // the tiles cover every voxel once): an agent whose next voxel, binned ahead by
    // the previous gather, is its current one takes the next stayer rank in slot
    // order, the others an atomic arrival rank in their next voxel.  The force
    // computation reads none of key, rank, stay or count, so each wave does this
    // after its agents' forces, beside the waves still computing theirs.
    // staged: the own voxels' slot ranges are in LDS (own_gs, own_pre) already.
    auto rank_tile = [&](bool staged) {
        if (!rank) return;
        const float inv_w = 1.0f / (float)W;
        for (uint32_t u = threadIdx.x; u < n_ownv; u += kBlock) {
            const uint32_t r = small_div(u, inv_w), c = u - r * W;
            const uint32_t g = voff + (T.a0 * n1 + T.b0 + r) * n2 + T.c0 + c;
            const uint32_t b = staged ? own_gs[u] : A.start[g];
            const uint32_t e = staged ? b + ((uint32_t)own_pre[u + 1] - (uint32_t)own_pre[u]) : A.start[g + 1];
            uint32_t n = 0;
            for (uint32_t s = b; s < e; s++) {
                const uint32_t k = A.key[s];
                if (k == g) A.rank[s] = n++;
                else A.rank[s] = kMoved | atomicAdd(&A.count[k], 1u);
            }
            A.stay[g] = n;
        }
    };

    if (fits) {
        // staged voxel v = (rb * NC + c - cl) * NP + pa; thread t holds v = 3t .. 3t + 2
        // (in 3-D the three planes of one (row, column)).  Their (plane, row, column)
        // come from one decomposition of 3t, then step by one voxel each.
        uint32_t cnt[3], gs[3], sum = 0;
        int pa_q[3], rb_q[3];
        uint32_t c_q[3];
        {
            const uint32_t v0 = 3 * threadIdx.x;
            uint32_t pa, rb, c;
            if constexpr (NP == 3) {
                pa = 0;
                rb = small_div(threadIdx.x, 1.0f / (float)G.NC);
                c = G.cl + (threadIdx.x - rb * G.NC);
            } else {
                pa = 0;
                rb = small_div(v0, 1.0f / (float)G.NC);
                c = G.cl + (v0 - rb * G.NC);
            }
#pragma unroll
            for (int q = 0; q < 3; q++) {
                pa_q[q] = (int)pa;
                rb_q[q] = (int)rb;
                c_q[q] = c;
                cnt[q] = 0;
                gs[q] = 0;
                const int a = (int)T.a0 + (int)pa - (int)PA_OWN, b = (int)T.b0 - 1 + (int)rb;
                if (v0 + q < n_vox && a >= 0 && a < (int)n0 && b >= 0 && b < (int)n1) {
                    const uint32_t g = voff + ((uint32_t)a * n1 + (uint32_t)b) * n2 + c;
                    gs[q] = A.start[g];
                    cnt[q] = A.start[g + 1] - gs[q];
                }
                sum += cnt[q];
                // next voxel in the staged order
                if constexpr (NP == 3) {
                    pa++;
                } else {
                    const bool wrap = c + 1 == G.ch;
                    c = wrap ? G.cl : c + 1;
                    rb = wrap ? rb + 1 : rb;
                }
            }
        }
        // the first kPreRecs records of each voxel are loaded before the scan, so
        // that their latency overlaps its barriers (mixed: 4 int4 records, 1M-agent
        // 3-D sorting force kernel 86.7 -> 84.6 us against 2; FP64 keeps 2 double4,
        // more would spill)
        constexpr uint32_t kPreRecs = MIXED ? 4 : 2;
        Rec pre[3][kPreRecs];
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (uint32_t k = 0; k < kPreRecs; k++)
                if (k < cnt[q]) pre[q][k] = rec[gs[q] + k];
        uint32_t total;
        uint32_t run = block_exclusive_scan(sum, total);
        const bool use_lds = total <= (uint32_t)CAP;  // uniform
        if (use_lds) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const uint32_t v = 3 * threadIdx.x + q;
                if (v < n_vox) {
                    vst[v] = (uint16_t)run;
                    const uint32_t pa = (uint32_t)pa_q[q], rb = (uint32_t)rb_q[q], c = c_q[q];
                    for (uint32_t k = 0; k < cnt[q]; k++) {
                        const Rec r = k < kPreRecs ? pre[q][k] : rec[gs[q] + k];
                        if constexpr (MIXED) lds[run + k] = to_region<D>(r, (int)pa, (int)rb, G);
                        else lds[run + k] = r;
                    }
                    if (P.sub_on && cnt[q]) {
                        const int a = (int)T.a0 + (int)pa - (int)PA_OWN, b = (int)T.b0 - 1 + (int)rb;
                        const uint16_t o = (uint16_t)subdomain_of(P, ((uint32_t)a * n1 + (uint32_t)b) * n2 + c);
                        for (uint32_t k = 0; k < cnt[q]; k++) lsub[run + k] = o;
                    }
                    // own voxel u = r * W + (c - c0): its count and first global slot
                    if (pa == PA_OWN && rb >= 1 && rb <= R && c >= T.c0 && c < T.c1) {
                        const uint32_t u = (rb - 1) * W + (c - T.c0);
                        own_pre[u + 1] = (uint16_t)cnt[q];
                        own_gs[u] = gs[q];
                    }
                    run += cnt[q];
                }
            }
            if (threadIdx.x == 0) {
                vst[n_vox] = (uint16_t)total;
                own_pre[0] = 0;
            }
            __syncthreads();
            {  // own_pre: exclusive prefix of the own voxel counts (n_ownv <= 256)
                const uint32_t u = threadIdx.x;
                const uint32_t x = u < n_ownv ? (uint32_t)own_pre[u + 1] : 0u;
                uint32_t tot;
                const uint32_t ex = block_exclusive_scan(x, tot);
                if (u < n_ownv) own_pre[u + 1] = (uint16_t)(ex + x);
            }
            __syncthreads();
            const uint32_t n_own = own_pre[n_ownv];
            LdsSource<D, MIXED> src{lds, vst, G, NP, n_vox, A.start, voff, n1, n2, lsub};
            const float inv_w = 1.0f / (float)W;
            // own agent t: own voxel u (largest u with own_pre[u] <= t), LDS index,
            // global slot, and the three ranges of its stream
            auto own_agent = [&](uint32_t t, uint32_t& self, Span& ov, Walk3& w) -> uint32_t {
                uint32_t lo = 0, hi = n_ownv;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if ((uint32_t)own_pre[mid] <= t) lo = mid;
                    else hi = mid;
                }
                const uint32_t k = t - (uint32_t)own_pre[lo];
                const uint32_t r = small_div(lo, inv_w), c = T.c0 + (lo - r * W);
                const uint32_t vo = src.vox(r + 1, c, PA_OWN);
                self = (uint32_t)vst[vo] + k;
                ov = Span{vst[vo], vst[vo + 1]};
                const uint32_t c_lo = c == 0 ? 0 : c - 1, c_hi = min(c + 2, n2);
                // stencil rows r, r + 1, r + 2 of the region (b - 1, b, b + 1; rows
                // outside the grid are staged empty)
                w.j = vst[src.vox(r, c_lo, 0)];
                w.je = vst[src.vox(r, c_hi, 0)];
                w.j1 = vst[src.vox(r + 1, c_lo, 0)];
                w.e1 = vst[src.vox(r + 1, c_hi, 0)];
                w.j2 = vst[src.vox(r + 2, c_lo, 0)];
                w.e2 = vst[src.vox(r + 2, c_hi, 0)];
                w.left = 2;
                return own_gs[lo] + k;
            };
            // Lanes take the own agents in descending order of candidate count, so
            // that a wave's lanes walk streams of similar length (a wave runs as
            // long as its longest stream).  Which lane computes an agent does not
            // change the agent's sums.
            // The sorted case looks each own agent up once and keeps the result in
            // LDS: self, own voxel range, the three stream ranges and the slot.
            const uint32_t gmask = (1u << lg) - 1u;  // lane within the agent's group
            const bool sorted = (n_own << lg) <= (uint32_t)kBlock;
            if (sorted) {
                if (threadIdx.x < kSortBins) bins[threadIdx.x] = 0;
                __syncthreads();
                uint32_t key = 0, kr = 0;
                if (threadIdx.x < n_own) {
                    uint32_t self;
                    Span ov;
                    Walk3 w;
                    const uint32_t s = own_agent(threadIdx.x, self, ov, w);
                    const uint32_t L = (w.je - w.j) + (w.e1 - w.j1) + (w.e2 - w.j2);
                    key = (uint32_t)(kSortBins - 1) - min(L >> 1, (uint32_t)(kSortBins - 1));
                    kr = atomicAdd(&bins[key], 1u);
                    OwnEntry& e = own_tab[threadIdx.x];
                    e.self = (uint16_t)self;
                    e.ovb = (uint16_t)ov.b;
                    e.ove = (uint16_t)ov.e;
                    e.r[0] = (uint16_t)w.j;
                    e.r[1] = (uint16_t)w.je;
                    e.r[2] = (uint16_t)w.j1;
                    e.r[3] = (uint16_t)w.e1;
                    e.r[4] = (uint16_t)w.j2;
                    e.r[5] = (uint16_t)w.e2;
                    e.slot = s;
                }
                __syncthreads();
                if (threadIdx.x < kSortBins) {  // exclusive scan of the class sizes (wave 0)
                    const uint32_t v = bins[threadIdx.x];
                    bins[threadIdx.x] = wave_incl_scan(v) - v;
                }
                __syncthreads();
                if (threadIdx.x < n_own) ord[bins[key] + kr] = (uint8_t)threadIdx.x;
                __syncthreads();
                if (threadIdx.x < (n_own << lg)) {
                    const OwnEntry e = own_tab[ord[threadIdx.x >> lg]];
                    Walk3 w;
                    w.j = e.r[0];
                    w.je = e.r[1];
                    w.j1 = e.r[2];
                    w.e1 = e.r[3];
                    w.j2 = e.r[4];
                    w.e2 = e.r[5];
                    w.left = 2;
                    agent_force<D, POT, MIXED, SENSE>(A, S, P, e.slot, e.self, lds[e.self], Span{e.ovb, e.ove}, w,
                                                      physical, symmetric, src, unit, MULTI ? threadIdx.x & gmask : 0u,
                                                      lg);
                }
                rank_tile(true);
                return;
            }
            for (uint32_t t = threadIdx.x; t < (n_own << lg); t += kBlock) {
                uint32_t self;
                Span ov;
                Walk3 w;
                const uint32_t s = own_agent(t >> lg, self, ov, w);
                agent_force<D, POT, MIXED, SENSE>(A, S, P, s, self, lds[self], ov, w, physical, symmetric, src, unit,
                                                  MULTI ? t & gmask : 0u, lg);
            }
            rank_tile(true);
            return;
        }
    }
    // global path: own agents row by row, candidates voxel by voxel
    GlobalSource<D, MIXED> src{rec, G};
    for (uint32_t r = 0; r < R; r++) {
        const uint32_t b = T.b0 + r, base = voff + (T.a0 * n1 + b) * n2;
        const uint32_t sb = A.start[base + T.c0], se = A.start[base + T.c1];
        for (uint32_t s = sb + threadIdx.x; s < se; s += kBlock) {
            const uint32_t c = S.vox[s] - base;
            WalkVox w;
            w.start = A.start;
            w.voff = voff;
            w.n1 = n1;
            w.n2 = n2;
            w.a_lo = T.a0 == 0 || NP == 1 ? T.a0 : T.a0 - 1;
            w.b_lo = b == 0 ? 0 : b - 1;
            w.c_lo = c == 0 ? 0 : c - 1;
            w.na = (NP == 1 ? T.a0 + 1 : min(T.a0 + 2, n0)) - w.a_lo;
            w.nb = min(b + 2, n1) - w.b_lo;
            w.ncol = min(c + 2, n2) - w.c_lo;
            w.rb = w.cc = w.pa = 0;
            const uint32_t g0 = w.gvox();
            w.j = A.start[g0];
            w.je = A.start[g0 + 1];
            const Span ov{A.start[base + c], A.start[base + c + 1]};
            WalkVox wm = w;  // the own record: its voxel (a0, b, c) in region coordinates
            wm.pa = T.a0 - w.a_lo;
            wm.rb = b - w.b_lo;
            agent_force<D, POT, MIXED, SENSE>(A, S, P, s, s, src.record(s, wm), ov, w, physical, symmetric, src, unit);
        }
    }
    rank_tile(false);
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

// Ring bookkeeping for a step that starts after t completed steps.
struct Ring {
    uint32_t cap, wslot, n_read;
    uint32_t it[2];  // slots yielded by the iterator after the push
};

__device__ __forceinline__ Ring ring_at(uint64_t t, uint32_t cap, uint32_t newest_first) {
    Ring r;
    r.cap = cap;
    r.wslot = cap ? (uint32_t)(t % cap) : 0u;
    const uint32_t size = (uint32_t)min<uint64_t>(t + 1, cap);
    const uint32_t oldest = size < cap ? 0u : (uint32_t)((t + 1) % (cap ? cap : 1u));
    r.n_read = (cap == 2 && size == 2) ? 2u : 1u;
#pragma unroll
    for (int i = 0; i < 2; i++)
        r.it[i] = cap == 0 ? 0u
                           : (newest_first ? (oldest + size - 1 - i + 2 * cap) % cap : (oldest + i) % cap);
    return r;
}

// The ring of a point agent whose first push is at step b (0, or the step after
// it was inserted by a division, whose AuxStorage starts empty): the entries are
// the pushes of steps max(b, t + 1 - cap) .. t (ring_at(t) when b == 0).  The
// capacity is 0, 1 or 2, so the slot arithmetic is a parity.
__device__ __forceinline__ Ring ring_agent(uint64_t t, uint64_t b, uint32_t cap, uint32_t newest_first) {
    Ring r;
    r.cap = cap;
    r.wslot = cap == 2 ? (uint32_t)(t & 1u) : 0u;
    const uint32_t size = (uint32_t)min<uint64_t>(t - b + 1, cap);
    const uint32_t oldest = cap == 2 ? (uint32_t)((t + 1 - size) & 1u) : 0u;
    r.n_read = (cap == 2 && size == 2) ? 2u : 1u;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const uint32_t q = newest_first ? oldest + size + 1u - (uint32_t)i : oldest + (uint32_t)i;
        r.it[i] = cap == 2 ? (q & 1u) : 0u;
    }
    return r;
}

// Standard normals of this agent and step (buffer or Philox).
template <int D>
__device__ __forceinline__ bool point_noise(const StepArgs& sa, uint32_t id, uint64_t t, double* z) {
    if (sa.noise_mode == CRM_NOISE_BUFFER && sa.noise) {
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = sa.noise[(uint64_t)id * D + k];
        return true;
    }
    if (sa.noise_mode == CRM_NOISE_PHILOX) {
        double g[4];
        gaussian_pair(sa.seed, id, t, 0, g[0], g[1]);
        if (D > 2) gaussian_pair(sa.seed, id, t, 1, g[2], g[3]);
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = g[k];
        return true;
    }
#pragma unroll
    for (int k = 0; k < D; k++) z[k] = 0.0;
    return false;
}

// One chili mechanics update of a point agent (solvers.rs:23-55,156-347 via
// local_mechanics_update, update_mechanics.rs:530-559) followed by
// apply_boundary (cartesian_cuboid_n.rs:594-631).  hist(i, dv) reads the value
// of the ring slot it[i] (never the slot written this step).  With
// WITH_VEL == false only the position is produced (the velocity, increments
// and, for NewtonDamped / Langevin, the noise do not influence it).
template <int D, int MECH, bool WITH_VEL, typename Hist>
__device__ __forceinline__ bool point_step(const InstParams& P, const StepArgs& sa, const Ring& rg, uint64_t t,
                                           uint32_t id, const double* x, const double* v, const double* f,
                                           Hist hist, double* nx, double* nv, double* dx, double* dv) {
    double z[D];
    const bool need_noise = WITH_VEL || MECH == CRM_MECHANICS_BROWNIAN;
    bool have_noise = false;
    if (need_noise) have_noise = point_noise<D>(sa, id, t, z);
    else {
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = 0.0;
    }
    double dxr[D], dvr[D];
    point_increment<MECH, D>(P, x, v, f, z, have_noise, sa.dt, sa.sqrt_dt, dx, dv, dxr, dvr);
    if (rg.cap == 0) {  // euler, solvers.rs:302-310
#pragma unroll
        for (int k = 0; k < D; k++) {
            nx[k] = (dx[k] * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) nv[k] = (dv[k] * sa.dt + v[k]) * 1.0 + dvr[k] * sa.dt;
        }
    } else if (rg.n_read == 2) {  // adams_bashforth_3, solvers.rs:312-329
        const double f0 = 23.0 / 12.0, f1 = (-16.0) / 12.0, f2 = 5.0 / 12.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double h0 = rg.it[0] == rg.wslot ? dx[k] : hist(0, k, false);
            const double h1 = rg.it[1] == rg.wslot ? dx[k] : hist(1, k, false);
            const double a = (dx[k] * f0 + (h0 * f1 + h1 * f2)) * sa.dt + x[k];
            nx[k] = a * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) {
                const double g0 = rg.it[0] == rg.wslot ? dv[k] : hist(0, k, true);
                const double g1 = rg.it[1] == rg.wslot ? dv[k] : hist(1, k, true);
                const double b = (dv[k] * f0 + (g0 * f1 + g1 * f2)) * sa.dt + v[k];
                nv[k] = b * 1.0 + dvr[k] * sa.dt;
            }
        }
    } else {  // adams_bashforth_2, solvers.rs:331-347
        const double g0c = 3.0 / 2.0, g1c = (-1.0) / 2.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double h0 = rg.it[0] == rg.wslot ? dx[k] : hist(0, k, false);
            nx[k] = ((dx[k] * g0c + h0 * g1c) * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) {
                const double q0 = rg.it[0] == rg.wslot ? dv[k] : hist(0, k, true);
                nv[k] = ((dv[k] * g0c + q0 * g1c) * sa.dt + v[k]) * 1.0 + dvr[k] * sa.dt;
            }
        }
    }
    if (WITH_VEL && MECH == CRM_MECHANICS_BROWNIAN) {
#pragma unroll
        for (int k = 0; k < D; k++) nv[k] = v[k];  // Brownian::set_velocity is a no-op
    }
    bool out = false;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        if (nx[k] < P.min[ax]) {
            nx[k] = 2.0 * P.min[ax] - nx[k];
            if (WITH_VEL) nv[k] = fabs(nv[k]);
        }
        if (nx[k] > P.max[ax]) {
            nx[k] = 2.0 * P.max[ax] - nx[k];
            if (WITH_VEL) nv[k] = -fabs(nv[k]);
        }
        out |= (nx[k] < P.min[ax]) || (nx[k] > P.max[ax]);
    }
    return out;
}

// Next voxel of every agent (get_index_of, cartesian_cuboid_n.rs:469-491) from
// the position it will have after this step (k_rank_ahead ranks it).  A
// position whose index has no voxel keeps the old one (update_mechanics.rs:430-453).
template <int D, int MECH>
__global__ __launch_bounds__(kBlock) void k_key_point(Arrays A, StateBuf S, StepArgs sa) {
    if (halted(A.ctrl, A.ctrl->step)) return;
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    const uint64_t t = A.ctrl->step;
    const uint32_t id = S.id[s];
    const Ring rg = ring_agent(t, S.birth ? S.birth[s] : 0u, sa.cap, sa.newest_first);
    double x[D], v[D], f[D], nx[D], nv[D], dx[D], dv[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        x[k] = S.pos[k * N + s];
        v[k] = MECH == CRM_MECHANICS_BROWNIAN ? 0.0 : S.vel[k * N + s];
        f[k] = MECH != CRM_MECHANICS_BROWNIAN ? 0.0 : A.forcef ? (double)A.forcef[k * N + s] : A.force[k * N + s];
    }
    auto hist = [&](int i, int k, bool) { return S.hdx[((uint64_t)rg.it[i] * D + k) * N + s]; };
    const bool out = point_step<D, MECH, false>(P, sa, rg, t, id, x, v, f, hist, nx, nv, dx, dv);
    if (out) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, t, id);
    uint32_t idx[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        ok = ok && axis_index_fast(nx[k], P.min[ax], P.dx[ax], P.inv_dx[ax], P.n[ax], idx[ax]);
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


// Each thread scans kScanItems consecutive voxels of count + stay into their
// offsets inside the scan tile (start_local) and clears count for the next step's
// ranks; whole groups move as two 16-byte loads per array and two 16-byte stores
// (count, stay and start_local are 256-byte aligned and a group starts at a
// multiple of 8 voxels).  bsum[tile] = the tile's sum.
__global__ __launch_bounds__(kBlock) void k_scan_local(uint32_t* __restrict__ count,
                                                        const uint32_t* __restrict__ stay,
                                                        uint32_t* __restrict__ start_local, uint32_t* __restrict__ bsum,
                                                        uint32_t Vt, const Control* ctrl) {
    if (halted(ctrl, ctrl->step)) return;
    static_assert(kScanItems == 8, "two uint4 per group");
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    const bool whole = base + kScanItems <= Vt;
    uint32_t vals[kScanItems];
    if (whole) {
        uint4* c4 = reinterpret_cast<uint4*>(count + base);
        const uint4* s4 = reinterpret_cast<const uint4*>(stay + base);
        const uint4 c0 = c4[0], c1 = c4[1], s0 = s4[0], s1 = s4[1];
        vals[0] = c0.x + s0.x;
        vals[1] = c0.y + s0.y;
        vals[2] = c0.z + s0.z;
        vals[3] = c0.w + s0.w;
        vals[4] = c1.x + s1.x;
        vals[5] = c1.y + s1.y;
        vals[6] = c1.z + s1.z;
        vals[7] = c1.w + s1.w;
        c4[0] = make_uint4(0u, 0u, 0u, 0u);
        c4[1] = make_uint4(0u, 0u, 0u, 0u);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            const uint64_t q = base + i;
            vals[i] = q < Vt ? count[q] + stay[q] : 0u;
            if (q < Vt) count[q] = 0u;
        }
    }
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) sum += vals[i];
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total);
    uint32_t out[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        out[i] = run;
        run += vals[i];
    }
    if (whole) {
        uint4* o4 = reinterpret_cast<uint4*>(start_local + base);
        o4[0] = make_uint4(out[0], out[1], out[2], out[3]);
        o4[1] = make_uint4(out[4], out[5], out[6], out[7]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; i++)
            if (base + i < Vt) start_local[base + i] = out[i];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place (the tile offsets);
// start[Vt] = total.  advance != 0 also completes the step counter (every kernel
// that reads it for this step has run; the gather reads the new value).
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t* __restrict__ bsum, uint32_t n_tiles,
                                                        uint32_t* __restrict__ start, uint32_t Vt, Control* ctrl,
                                                        uint32_t advance) {
    const bool skip = halted(ctrl, ctrl->step);
    __syncthreads();  // every thread has read the control block before thread 0 writes it
    if (threadIdx.x == 0) ctrl->skip = skip ? 1u : 0u;
    if (skip) return;
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

// Thread g writes the final start of voxel g (tile-local offset + tile offset) and
// places agent g in the slot list of its next voxel: a stayer at its rank, an
// arrival (kMoved) after the voxel's stayers in atomic order.  Launched over
// max(N, Vt) threads.
__global__ __launch_bounds__(kBlock) void k_scatter(Arrays A) {
    if (A.ctrl->skip) return;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g < A.Vt) A.start[g] = A.start_local[g] + A.tile_off[g / kScanTile];
    if (g >= A.N) return;
    const uint32_t k = A.key[g], r = A.rank[g];
    const uint32_t pos = (r & kMoved) ? A.stay[k] + (r & ~kMoved) : r;
    A.perm[A.start_local[k] + A.tile_off[k / kScanTile] + pos] = (uint32_t)g;
}

// Small systems (every voxel in one scan tile, at most kSmallSortAgents agents):
// k_scan_local, k_scan_tiles and k_scatter in one workgroup, with barriers in place
// of the launches between them; the voxel starts stay in LDS for the scatter.  The
// same skip flag, count reset, starts, start[Vt], step advance and slot lists.
constexpr uint32_t kSmallSortAgents = 16384;
__global__ __launch_bounds__(kBlock) void k_sort_small(Arrays A, uint32_t advance) {
    static_assert(kScanItems == 8, "eight voxels per thread");
    __shared__ uint32_t s_start[kScanTile];
    const bool skip = halted(A.ctrl, A.ctrl->step);
    __syncthreads();  // every thread has read the control block before thread 0 writes it
    if (threadIdx.x == 0) A.ctrl->skip = skip ? 1u : 0u;
    if (skip) return;
    const uint32_t Vt = A.Vt, base = threadIdx.x * kScanItems;
    uint32_t vals[kScanItems], sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const uint32_t q = base + i;
        vals[i] = q < Vt ? A.count[q] + A.stay[q] : 0u;
        if (q < Vt) A.count[q] = 0u;
        sum += vals[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const uint32_t q = base + i;
        if (q < Vt) {
            s_start[q] = run;
            A.start[q] = run;
        }
        run += vals[i];
    }
    if (threadIdx.x == 0) {
        A.start[Vt] = total;
        if (advance) A.ctrl->step += 1;
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < (uint32_t)A.N; g += kBlock) {
        const uint32_t k = A.key[g], r = A.rank[g];
        const uint32_t pos = (r & kMoved) ? A.stay[k] + (r & ~kMoved) : r;
        A.perm[s_start[k] + pos] = g;
    }
}

// Orders each voxel's arrivals by previous slot (atomic arrival order is not
// deterministic).  The resulting list is the reference's: the cells that stay keep
// their order and the movers are appended in order of old voxel, then old position
// (sort_cells_in_voxels_step_1/2, update_mechanics.rs:418-513).  Set-up sorts rank
// every agent atomically with stay = 0 and get the whole voxel in order of previous
// slot.  The point gather of a step orders the arrivals itself (arrival_slot), so
// the step's sort launches this kernel only for rods.
__global__ __launch_bounds__(kBlock) void k_fixup(Arrays A) {
    if (A.ctrl->skip) return;
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= A.Vt) return;
    // the stayers hold [b, b + stay) in their old order; the arrivals follow, put in
    // order of previous slot, i.e. of (old voxel, old position)
    const uint32_t m = A.start[v] + A.stay[v], e = A.start[v + 1];
    for (uint32_t i = m + 1; i < e; i++) {
        uint32_t x = A.perm[i];
        uint32_t j = i;
        while (j > m) {
            uint32_t y = A.perm[j - 1];
            if (y <= x) break;
            A.perm[j] = y;
            j--;
        }
        A.perm[j] = x;
    }
}

// Slot of the agent from previous slot o that k_scatter put at slot i of voxel key:
// an arrival (kMoved rank) belongs after the voxel's stayers at the number of the
// voxel's arrivals with a smaller previous slot, the order k_fixup gives.
__device__ __forceinline__ uint32_t arrival_slot(const Arrays& A, uint32_t i, uint32_t o, uint32_t key) {
    if (!(A.rank[o] & kMoved)) return i;
    const uint32_t m = A.start[key] + A.stay[key], e = A.start[key + 1];
    if (e - m < 2u) return i;
    uint32_t r = 0;
    for (uint32_t j = m; j < e; j++) r += A.perm[j] < o ? 1u : 0u;
    return m + r;
}

// Force record of slot s: absolute double coordinates (FP64), or 30-bit fixed
// point relative to the voxel origin (mixed; the sign bit of w flags an agent
// outside its voxel's fixed-point range, which sends its pairs to the FP64 path).
template <int D, bool MIXED>
__device__ __forceinline__ void write_record(const Arrays& A, const InstParams& P, uint64_t s, uint32_t key,
                                             const double* p, int32_t sp) {
    if constexpr (MIXED) {
        const uint32_t lv = key - P.voxel_offset;
        uint32_t idx[3];
        const uint32_t row = udiv_by_inv(lv, P.n[2], P.inv_n2);  // lv / n2
        idx[2] = lv - row * P.n[2];
        idx[0] = udiv_by_inv(lv, P.n[2] * P.n[1], P.inv_n12);
        idx[1] = row - idx[0] * P.n[1];
        int q[3] = {0, 0, 0};
        uint32_t outside = 0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const int ax = k + 3 - D;
            const double origin = P.min[ax] + (double)idx[ax] * P.dx[ax];
            const double tq = floor((p[k] - origin) * P.qinv[ax]);
            const bool ok = tq >= 0.0 && tq < (double)kQOne;  // false for NaN
            q[k] = ok ? (int)tq : (tq < 0.0 ? 0 : kQOne - 1);
            outside |= ok ? 0u : 1u;
        }
        const uint32_t w = ((uint32_t)sp & kSpeciesMask) | ((idx[2] & kColMask) << kColShift) | (outside << 31);
        A.recq[s] = make_int4(q[0], q[1], D == 3 ? q[2] : 0, (int)w);
    } else {
        A.rec64[s] = make_double4(p[0], p[1], D == 3 ? p[D - 1] : 0.0, (double)sp);
    }
}

// Next voxel of the agent in slot s for the step that starts after t1 steps,
// from its state at that point: position x, velocity v and the ring slots r0 /
// r1 of position increments.  For NewtonDamped and Langevin the position update
// depends on neither the force nor the noise, so the re-binning of step t1 is
// known once step t1 - 1 has been integrated (k_key_point computes the same
// position).  A position whose index has no voxel keeps `vox`.
template <int D, int MECH>
__device__ __forceinline__ void bin_ahead(const Arrays& A, const InstParams& P, const StepArgs& sa, uint64_t t1,
                                          uint32_t id, uint64_t s, const double* x, const double* v,
                                          const double* r0, const double* r1, uint32_t vox, uint64_t birth) {
    const Ring rg = ring_agent(t1, birth, sa.cap, sa.newest_first);
    double f[D], nx[D], nv[D], dx[D], dv[D];
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = 0.0;
    auto hist = [&](int i, int k, bool) { return rg.it[i] ? r1[k] : r0[k]; };
    point_step<D, MECH, false>(P, sa, rg, t1, id, x, v, f, hist, nx, nv, dx, dv);
    uint32_t idx[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        ok = ok && axis_index_fast(nx[k], P.min[ax], P.dx[ax], P.inv_dx[ax], P.n[ax], idx[ax]);
    }
    uint32_t key = vox;
    if (ok) key = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2];
    else atomicAdd((unsigned long long*)&A.ctrl->out_of_grid, 1ull);
    A.key_next[s] = key;
}

// Ranks of the next voxels (binned ahead by the gather, or by a key kernel after
// the forces), with an atomic only for agents that change voxel.  Thread per
// voxel v of the current order: the agents of v that stay get consecutive ranks
// in slot order and their number goes to stay[v]; the others take an atomic rank
// among the arrivals of their next voxel (count), flagged with kMoved, which the
// scatter places after that voxel's stayers; k_fixup puts the arrivals in order.
__global__ __launch_bounds__(kBlock) void k_rank_ahead(Arrays A) {
    if (halted(A.ctrl, A.ctrl->step)) return;
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= A.Vt) return;
    const uint32_t b = A.start[v], e = A.start[v + 1];
    uint32_t n = 0;
    for (uint32_t s = b; s < e; s++) {
        const uint32_t k = A.key[s];
        if (k == (uint32_t)v) A.rank[s] = n++;
        else A.rank[s] = kMoved | atomicAdd(&A.count[k], 1u);
    }
    A.stay[v] = n;
}

// Permutes the state into sorted order and writes the force records, without
// integrating (initial placement).  The ring holds min(step, cap) increments.
// With AHEAD the next step's voxels are binned as well.
template <int D, bool MIXED, bool AHEAD, int MECH>
__global__ __launch_bounds__(kBlock) void k_gather_point(Arrays A, StateBuf src, StateBuf dst, StepArgs sa) {
    if (A.ctrl->skip) return;
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint64_t s = w.first + threadIdx.x;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    const uint32_t o = A.perm[s];
    const uint64_t t = A.ctrl->step;
    const uint64_t birth = src.birth ? src.birth[o] : 0u;
    // every slot when agents hold rings of different fill (inserted daughters)
    const uint32_t n_hist = src.birth ? sa.cap : (uint32_t)min<uint64_t>(t, sa.cap);
    double p[D], v[D], r0[D], r1[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        p[k] = src.pos[k * N + o];
        v[k] = src.vel[k * N + o];
        dst.pos[k * N + s] = p[k];
        dst.vel[k * N + s] = v[k];
        r0[k] = r1[k] = 0.0;
    }
    for (uint32_t h = 0; h < n_hist; h++) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double hx = src.hdx[((uint64_t)h * D + k) * N + o];
            dst.hdx[((uint64_t)h * D + k) * N + s] = hx;
            dst.hdv[((uint64_t)h * D + k) * N + s] = src.hdv[((uint64_t)h * D + k) * N + o];
            if (h == 0) r0[k] = hx;
            else r1[k] = hx;
        }
    }
    const int32_t sp = src.species[o];
    const uint32_t id = src.id[o];
    dst.species[s] = sp;
    dst.id[s] = id;
    const uint32_t key = A.key[o];
    dst.vox[s] = key;
    write_record<D, MIXED>(A, P, s, key, p, sp);
    if (dst.birth) dst.birth[s] = (uint32_t)birth;
    if constexpr (AHEAD) bin_ahead<D, MECH>(A, P, sa, t, id, s, p, v, r0, r1, key, birth);
}

// Gather + integrate: thread i moves the agent from slot o = perm[i] to its sorted
// slot s (i itself, or arrival_slot's place among the voxel's arrivals); its
// update (the same point_step as in k_key_point, so the same position) is computed
// here and the new state written straight to slot s, together with the force
// record for the next step.  k_scan_tiles has already advanced the step counter,
// so this step started after step - 1 steps.
template <int D, bool MIXED, int MECH>
__global__ __launch_bounds__(kBlock) void k_gather_integrate_point(Arrays A, StateBuf src, StateBuf dst,
                                                                     StepArgs sa) {
    if (A.ctrl->skip) return;
    const BlockWork w = A.work[blockIdx.x];
    if (threadIdx.x >= w.count) return;
    const uint32_t i = (uint32_t)(w.first + threadIdx.x);
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    const uint64_t t = A.ctrl->step - 1;
    const uint32_t o = A.perm[i];
    const uint32_t key = A.key[o];
    const uint64_t s = arrival_slot(A, i, o, key);
    const uint32_t id = src.id[o];
    const uint64_t birth = src.birth ? src.birth[o] : 0u;
    const Ring rg = ring_agent(t, birth, sa.cap, sa.newest_first);
    double x[D], v[D], f[D], nx[D], nv[D], dx[D], dv[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        x[k] = __builtin_nontemporal_load(&src.pos[k * N + o]);
        v[k] = __builtin_nontemporal_load(&src.vel[k * N + o]);
        if constexpr (MIXED) f[k] = (double)__builtin_nontemporal_load(&A.forcef[k * N + o]);
        else f[k] = __builtin_nontemporal_load(&A.force[k * N + o]);
    }
    // the ring before this step (min(t - birth, cap) entries: with cap <= 2 the
    // slot not overwritten now, once filled), in registers, except the slot this
    // step overwrites: point_step reads the new increment in its place
    const uint32_t n_old = (uint32_t)min<uint64_t>(t - birth, rg.cap);
    double hx[2][D], hv[2][D];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bool live = (uint32_t)h < rg.cap && (uint32_t)h != rg.wslot && n_old > 0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const uint64_t e = ((uint64_t)h * D + k) * N + o;
            hx[h][k] = live ? __builtin_nontemporal_load(&src.hdx[e]) : 0.0;
            hv[h][k] = live ? __builtin_nontemporal_load(&src.hdv[e]) : 0.0;
        }
    }
    auto hist = [&](int i, int k, bool is_dv) {
        const uint32_t sl = rg.it[i];
        return is_dv ? (sl ? hv[1][k] : hv[0][k]) : (sl ? hx[1][k] : hx[0][k]);
    };
    const bool out = point_step<D, MECH, true>(P, sa, rg, t, id, x, v, f, hist, nx, nv, dx, dv);
    if (out) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, t, id);
#pragma unroll
    for (int k = 0; k < D; k++) {
        __builtin_nontemporal_store(nx[k], &dst.pos[k * N + s]);
        __builtin_nontemporal_store(nv[k], &dst.vel[k * N + s]);
    }
    // ring: the slot written this step gets (dx, dv), the other slot is carried over
    if (rg.cap > 0) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const bool keep = (uint32_t)h < rg.cap && ((uint32_t)h == rg.wslot || n_old > 0);
#pragma unroll
            for (int k = 0; k < D; k++) {
                if (h == (int)rg.wslot) {
                    hx[h][k] = dx[k];
                    hv[h][k] = dv[k];
                }
                if (keep) {
                    const uint64_t e = ((uint64_t)h * D + k) * N + s;
                    __builtin_nontemporal_store(hx[h][k], &dst.hdx[e]);
                    __builtin_nontemporal_store(hv[h][k], &dst.hdv[e]);
                }
            }
        }
    }
    const int32_t sp = src.species[o];
    dst.species[s] = sp;
    dst.id[s] = id;
    dst.vox[s] = key;
    write_record<D, MIXED>(A, P, s, key, nx, sp);
    if (dst.birth) dst.birth[s] = (uint32_t)birth;
    if constexpr (MECH != CRM_MECHANICS_BROWNIAN)
        bin_ahead<D, MECH>(A, P, sa, t + 1, id, s, nx, nv, hx[0], hx[1], key, birth);
}

}  // namespace crm
