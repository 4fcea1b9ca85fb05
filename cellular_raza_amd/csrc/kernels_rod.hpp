// kernels_rod.hpp -- gfx950 kernels of the mechanics step for bacterial rods.
//
// A rod is a chain of nv vertices (RodMechanics, bacterial_rods.rs:82-265); it is
// binned by its vertex centroid (bacterial_rods.rs:354-373,536-555).  One lane
// handles one (rod, vertex); the lanes of a rod are consecutive in a workgroup.
// Step as launched by the runtime:
//   k_force_rod              RodInteraction (bacterial_rods.rs:292-336) with the chili
//                            pair weights (update_mechanics.rs:178-220)
//   k_key_rod                next vertex positions, per-vertex reflection
//                            (bacterial_rods.rs:489-533), centroid voxel + rank
//   k_scan_local/tiles, k_scatter, k_fixup   as for point agents (slots are rods)
//   k_gather_integrate_rod   RodMechanics::calculate_increment (bacterial_rods.rs:157-240)
//                            + solver, written into the rod's new slot
// State arrays are the point arrays with K = nv * dim components per slot; the
// force kernel reads rods from an AoS copy (rec_rod[slot][K]) written by the gathers.
#pragma once

#include "kernels_point.hpp"

namespace crm {

struct RodArgs {
    double* rec;   // [N][K] AoS vertex positions (force records)
    double4* bnd;  // [N] bounding ball of each rod: vertex mean (x, y, z), radius
    float4* seg;   // [N][nv - 1] ball of each segment: midpoint minus the rod's vertex
                   // mean (x, y, z), half length with margin (infinite for a
                   // zero-length segment)
    uint32_t nv;   // vertices per rod
    uint32_t rpb;  // rods per workgroup
};

// Lane -> (rod slot, vertex).  Returns false for lanes past the block's rods.
__device__ __forceinline__ bool rod_lane(const BlockWork& w, uint32_t nv, uint64_t& s, uint32_t& v) {
    const uint32_t r = threadIdx.x / nv;
    v = threadIdx.x % nv;
    s = w.first + r;
    return r < w.count;
}

// ---------------------------------------------------------------------------
// Force kernel.  For own rod i, vertex v and every candidate rod j of the 3^D
// stencil of i's voxel:
//   own part  (calc_i(i, j), row v):  sum over segments s of j of the point force
//             between vertex v and its nearest point on segment s;
//   ext part  (calc_j(j, i), row v):  for every vertex u of j, the force on the
//             nearest point of i's segments (v-1, v) (weight t) and (v, v+1)
//             (weight 1 - t), in the reference's (u, s) order.
// Reference weights: same voxel 1/2 own + 1/2 ext; adjacent voxel 1/2 own
// (physical: same-voxel weights everywhere).
//
// A rod's segments lie inside the ball around its vertex mean that holds its
// vertices, and a segment inside the ball around its midpoint of half its
// length: a point farther than (radius + cutoff) from such a ball has every
// nearest point beyond the cutoff, where every term is exactly zero.  Pass 1
// marks the (candidate, segment) pairs that can be inside the cutoff, pass 2
// evaluates only those.  Own parts are summed per candidate in the reference's
// order; the ext parts follow the own parts of all candidates.
// ---------------------------------------------------------------------------
// Staging: the block's rods lie in at most two rows of voxels (own rows; the slots
// are in voxel order), and every candidate of a rod lies in one of the 3^(D-1)
// stencil rows of its own row, within the columns of the own row's rods +- 1.  Those
// rows' rods (records, bounding and segment balls, species) are copied to LDS, up
// to kRodLdsBytes, so that the candidate reads of the walks below come from LDS; a
// block with more own rows or more rods reads global memory (same order, same
// arithmetic).  36 KiB hold the 2-D 8-vertex rods' stencil rows of nearly every
// block (200,000 2-D rods: 32 KiB left too many blocks on global memory, 418 ->
// 409 us; 36 KiB 348 us; 40 KiB 347 us) and leave 4 workgroups per CU.
constexpr uint32_t kRodLdsBytes = 36 * 1024;
template <int D, int POT, bool MIXED>
__global__ __launch_bounds__(kBlock, 4) void k_force_rod(Arrays A, StateBuf S, RodArgs R, uint32_t physical,
                                                          uint32_t no_lds) {
    if (halted(A.ctrl, A.ctrl->step)) return;
    const BlockWork w = A.work[blockIdx.x];
    uint64_t s;
    uint32_t v;
    const bool active = rod_lane(w, R.nv, s, v);
    const InstParams& P = A.ip[w.inst];
    const uint32_t nv = R.nv, K = nv * D, nseg = nv - 1;
    __shared__ __align__(16) char lbuf[kRodLdsBytes];
    __shared__ uint32_t rg_slot[18], rg_cnt[18], rg_off[19];
    __shared__ uint32_t s_row[2], s_cmin[2], s_cmax[2], s_bad;
    const uint32_t n1 = P.n[1], n2 = P.n[2], voff = P.voxel_offset;
    const uint32_t per_rod = 8u * K + 32u + 16u * nseg + 4u;
    const uint32_t max_rods = (kRodLdsBytes - 64u) / per_rod;  // room for the sections' 16-byte alignment
    // LDS layout, each section 16-byte aligned: bnd [max_rods] double4, rec
    // [max_rods][K] doubles, seg [max_rods][nseg] float4, species [max_rods] int
    auto up16 = [](uint32_t b) { return (b + 15u) & ~15u; };
    const uint32_t o_rec = up16(32u * max_rods), o_seg = up16(o_rec + 8u * K * max_rods);
    const uint32_t o_sp = up16(o_seg + 16u * nseg * max_rods);
    double4* lbnd = reinterpret_cast<double4*>(lbuf);
    double* lrec = reinterpret_cast<double*>(lbuf + o_rec);
    float4* lseg = reinterpret_cast<float4*>(lbuf + o_seg);
    int* lsp = reinterpret_cast<int*>(lbuf + o_sp);
    if (threadIdx.x == 0) {
        s_row[0] = (S.vox[w.first] - voff) / n2;
        s_row[1] = (S.vox[w.first + w.count - 1] - voff) / n2;
        s_cmin[0] = s_cmin[1] = 0xffffffffu;
        s_cmax[0] = s_cmax[1] = 0u;
        s_bad = no_lds;
    }
    __syncthreads();
    if (threadIdx.x < w.count) {  // own rows and their column ranges
        const uint32_t lv = S.vox[w.first + threadIdx.x] - voff, row = lv / n2, c = lv - row * n2;
        const uint32_t q = row == s_row[0] ? 0u : (row == s_row[1] ? 1u : 2u);
        if (q == 2u) s_bad = 1u;
        else {
            atomicMin(&s_cmin[q], c);
            atomicMax(&s_cmax[q], c);
        }
    }
    __syncthreads();
    if (threadIdx.x < 18) {  // range (own row q, stencil row (da, db))
        const uint32_t q = threadIdx.x / 9u, t9 = threadIdx.x - q * 9u;
        const int da = (int)(t9 / 3u) - 1, db = (int)(t9 % 3u) - 1;
        uint32_t slot = 0, cnt = 0;
        if (q == 0 || s_row[1] != s_row[0]) {
            const uint32_t row = s_row[q], a = row / n1, b = row - a * n1;
            const int aa = (int)a + da, bb = (int)b + db;
            if (aa >= 0 && aa < (int)P.n[0] && bb >= 0 && bb < (int)n1) {
                const uint32_t clo = s_cmin[q] == 0 ? 0u : s_cmin[q] - 1u, chi = min(s_cmax[q] + 2u, n2);
                const uint32_t g = voff + ((uint32_t)aa * n1 + (uint32_t)bb) * n2;
                slot = A.start[g + clo];
                cnt = A.start[g + chi] - slot;
            }
        }
        rg_slot[threadIdx.x] = slot;
        rg_cnt[threadIdx.x] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        for (int r = 0; r < 18; r++) {
            rg_off[r] = o;
            o += rg_cnt[r];
        }
        rg_off[18] = o;
        if (o > max_rods) s_bad = 1u;
    }
    __syncthreads();
    const bool staged = s_bad == 0u;  // uniform
    if (staged) {
        const uint32_t total = rg_off[18];
        for (uint32_t e = threadIdx.x; e < total; e += kBlock) {  // thread per staged rod
            uint32_t r = 0;
            while (rg_off[r + 1] <= e) r++;
            const uint32_t j = rg_slot[r] + (e - rg_off[r]);
            lbnd[e] = R.bnd[j];
            lsp[e] = S.species[j];
            for (uint32_t k = 0; k < K; k++) lrec[e * K + k] = R.rec[(uint64_t)j * K + k];
            for (uint32_t q = 0; q < nseg; q++) lseg[e * nseg + q] = R.seg[(uint64_t)j * nseg + q];
        }
    }
    __syncthreads();
    if (!active) return;
    // this lane's own row among the block's (0 or 1)
    const uint32_t orow = ((S.vox[s] - voff) / n2) == s_row[0] ? 0u : 1u;
    auto work = [&](auto stg) {
        constexpr bool STG = decltype(stg)::value;
        auto bnd_of = [&](uint32_t j, int32_t base) -> double4 {
            if constexpr (STG) return lbnd[(int32_t)j + base];
            else return R.bnd[j];
        };
        auto seg_of = [&](uint32_t j, int32_t base) -> const float4* {
            if constexpr (STG) return lseg + (uint32_t)((int32_t)j + base) * nseg;
            else return R.seg + (uint64_t)j * nseg;
        };
        auto rec_of = [&](uint32_t j, int32_t base) -> const double* {
            if constexpr (STG) return lrec + (uint32_t)((int32_t)j + base) * K;
            else return R.rec + (uint64_t)j * K;
        };
        auto sp_of = [&](uint32_t j, int32_t base) -> int {
            if constexpr (STG) return lsp[(int32_t)j + base];
            else return S.species[j];
        };
    const PotParams<double>& PP = P.pd;
    const double* own = R.rec + s * K;
    double pv[D];
#pragma unroll
    for (int k = 0; k < D; k++) pv[k] = own[v * D + k];
    const uint32_t lv = S.vox[s] - P.voxel_offset;
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
    const int sp_i = S.species[s];
    double acc[D];
#pragma unroll
    for (int k = 0; k < D; k++) acc[k] = 0.0;
    bool calc_err = false;
    // Reach tests in float from double differences: a ball of radius r around
    // centre c is out of reach of point p when |p - c| > r + cutoff.  The margin
    // (1e-5 absolute + 1e-6 relative) is far above the float rounding of
    // distances within a stencil.
    const double rcd = cut_radius<POT, double>(PP);
    float reach = (float)rcd * (1.0f + 1e-6f) + 1e-5f;
    if (!(reach >= 0.0f)) reach = __builtin_inff();
    auto out_of_reach = [&](const double* p, const double* c, float radius) {
        float d2 = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const float d = (float)(p[k] - c[k]);
            d2 = fmaf(d, d, d2);
        }
        const float lim = radius + reach;
        return d2 > lim * lim;
    };
    // segment (a, b): midpoint m and half length h (+ margin); zero-length
    // segments get an infinite radius, so they are always evaluated (their
    // nearest point is undefined, as in the reference)
    auto seg_ball = [&](const double* a, const double* b, double* m) {
        float h2 = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            m[k] = (a[k] + b[k]) * 0.5;
            const float d = (float)(b[k] - a[k]);
            h2 = fmaf(d, d, h2);
        }
        bool zero = true;
#pragma unroll
        for (int k = 0; k < D; k++) zero = zero && a[k] == b[k];
        return zero ? __builtin_inff() : 0.5f * sqrtf(h2) * (1.0f + 1e-6f) + 1e-6f;
    };
    const double4 bi = R.bnd[s];
    const double ci[3] = {bi.x, bi.y, bi.z};
    const float ri = (float)bi.w * (1.0f + 1e-6f);
    // the own voxel's rods (same-voxel pairs)
    const uint32_t own_row = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2];
    const uint32_t cb = A.start[own_row + idx[2]], ce = A.start[own_row + idx[2] + 1];
    // One vertex-segment evaluation: the force on point p from its nearest point
    // on segment (a, b), and the segment parameter t of that point.  Mixed
    // precision: the segment relative to p in float, the potential in float; a
    // pair whose squared distance lies within 2e-6 (relative) of the cutoff
    // radius's, a distance below a tenth of p's distance to the segment ends, or
    // a zero-length segment is evaluated in FP64.
    const float rc2f = (float)(rcd * rcd), bandf = 2e-6f * rc2f;
    auto pair = [&](const double* p, const double* a, const double* b, bool same, double strength, double* f,
                    double& t) {
        if constexpr (MIXED) {
            float av[D], ab[D], num = 0.0f, den = 0.0f, span = 0.0f;
#pragma unroll
            for (int k = 0; k < D; k++) {
                av[k] = (float)(a[k] - p[k]);
                const float bk = (float)(b[k] - p[k]);
                ab[k] = bk - av[k];
                num = num - av[k] * ab[k];
                den = den + ab[k] * ab[k];
                span = span + av[k] * av[k] + bk * bk;
            }
            // (v_rcp_f32: 1 ulp on a parameter clamped to [0, 1]; den == 0 gives a
            // non-finite tt and the FP64 path below, as a division would)
            float tt = num * __builtin_amdgcn_rcpf(den);
            tt = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
            float z[D], d2 = 0.0f;
#pragma unroll
            for (int k = 0; k < D; k++) {
                z[k] = -(av[k] + ab[k] * tt);
                d2 = d2 + z[k] * z[k];
            }
            // (a distance below a tenth of the endpoints' would carry their float
            // rounding into the direction at more than ~1e-6)
            if (den > 0.0f && d2 > 1e-2f * span && fabsf(d2 - rc2f) > bandf) {
                float ff[D];
                const bool e = point_force<POT, float, D>(z, same, (float)strength, P.pf, -1, ff);
#pragma unroll
                for (int k = 0; k < D; k++) f[k] = (double)ff[k];
                t = (double)tt;
                return e;
            }
        }
        double x[D], z[D];
        nearest_point<double, D>(p, a, b, x, t);
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = p[k] - x[k];
        return point_force<POT, double, D>(z, same, strength, PP, -1, f);
    };
    // own segments (v - 1, v) and (v, v + 1): balls for the ext part
    double m0[D], m1[D];
    float h0 = -1.0f, h1 = -1.0f;  // negative: no such segment
    if (v >= 1) h0 = seg_ball(own + (v - 1) * D, own + v * D, m0);
    if (v + 1 < nv) h1 = seg_ball(own + v * D, own + (v + 1) * D, m1);

    struct RodWalk {
        uint32_t ra, rb, j, je;
        int32_t base;  // staged: LDS index of slot j = j + base in this stencil row
    };
    // the staged range of stencil row (ra, rb) of this lane's own row
    auto row_base = [&](uint32_t ra, uint32_t rb) -> int32_t {
        if constexpr (!STG) return 0;
        const uint32_t r = orow * 9u + (ra + 1u - idx[0]) * 3u + (rb + 1u - idx[1]);
        return (int32_t)rg_off[r] - (int32_t)rg_slot[r];
    };
    // candidates: stencil rows (ra, rb) ascending, each the rods of columns [lo2, hi2)
    auto settle = [&](RodWalk& wk) {
        while (wk.j >= wk.je) {
            const uint32_t over = wk.j - wk.je;
            if (++wk.rb == hi[1]) {
                wk.rb = lo[1];
                if (++wk.ra == hi[0]) return false;
            }
            const uint32_t row = P.voxel_offset + (wk.ra * P.n[1] + wk.rb) * P.n[2];
            wk.j = A.start[row + lo[2]] + over;
            wk.je = A.start[row + hi[2]];
            wk.base = row_base(wk.ra, wk.rb);
        }
        return true;
    };
    auto first = [&](RodWalk& wk) {
        wk.ra = lo[0];
        wk.rb = lo[1];
        const uint32_t row = P.voxel_offset + (wk.ra * P.n[1] + wk.rb) * P.n[2];
        wk.j = A.start[row + lo[2]];
        wk.je = A.start[row + hi[2]];
        wk.base = row_base(wk.ra, wk.rb);
        return settle(wk);
    };
    auto strengths = [&](uint32_t j, int32_t base, bool& same, double& s_own, double& s_ext) {
        const int sp_j = sp_of(j, base);
        same = sp_i == sp_j;
        s_own = s_ext = PP.strength;
        if constexpr (POT == CRM_POTENTIAL_MORSE_SPECIES) {
            s_own = PP.strength * PP.species_matrix[sp_i * CRM_MAX_SPECIES + sp_j];
            s_ext = PP.strength * PP.species_matrix[sp_j * CRM_MAX_SPECIES + sp_i];
        }
    };

    // Own part, calc_i(i, j) row v: per candidate rod, the point force between
    // vertex v and its nearest point on each segment, summed over segments.
    // Chunks of `per` candidates; pass 1 sets bit (c * nseg + q) for each segment
    // q of candidate c that can reach v, pass 2 evaluates the set bits in order.
    const uint32_t per = 64u / nseg < 32u ? 64u / nseg : 32u;
    RodWalk cur;
    bool live = POT != CRM_POTENTIAL_NONE && first(cur);
    while (live) {
        const RodWalk c0 = cur;
        uint64_t mo = 0;
        for (uint32_t c = 0; c < per; c++) {
            const uint32_t j = cur.j;
            if (j != s) {
                // v relative to j's vertex mean in float, then to each segment ball
                // (their offsets from the mean are small, so the float differences
                // stay far inside the reach margin)
                const double4 bj = bnd_of(j, cur.base);
                const double cj[3] = {bj.x, bj.y, bj.z};
                float dc[D], d2 = 0.0f;
#pragma unroll
                for (int k = 0; k < D; k++) {
                    dc[k] = (float)(pv[k] - cj[k]);
                    d2 = fmaf(dc[k], dc[k], d2);
                }
                const float lim = (float)bj.w * (1.0f + 1e-6f) + reach;
                if (!(d2 > lim * lim)) {
                    const float4* sg = seg_of(j, cur.base);
                    for (uint32_t q = 0; q < nseg; q++) {
                        const float4 g = sg[q];
                        const float go[3] = {g.x, g.y, g.z};
                        float e2 = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; k++) {
                            const float e = dc[k] - go[k];
                            e2 = fmaf(e, e, e2);
                        }
                        const float l = g.w + reach;
                        if (!(e2 > l * l)) mo |= 1ull << (c * nseg + q);
                    }
                }
            }
            ++cur.j;
            live = settle(cur);
            if (!live) break;
        }
        // pass 2: bits in order; a candidate's segment sum goes to acc (x 1/2)
        // when the next candidate starts
        RodWalk wk = c0;
        uint32_t pos = 0, cand = 0xffffffffu;
        const double* ext = nullptr;
        bool same = false;
        double s_own = 0.0, s_ext = 0.0;
        double fo_sum[D];
#pragma unroll
        for (int k = 0; k < D; k++) fo_sum[k] = 0.0;
        while (mo) {
            const uint32_t bit = (uint32_t)__builtin_ctzll(mo);
            mo &= mo - 1;
            const uint32_t c = bit / nseg, q = bit - c * nseg;
            if (c != cand) {
                if (cand != 0xffffffffu) {
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        acc[k] = acc[k] + fo_sum[k] * 0.5;
                        fo_sum[k] = 0.0;
                    }
                }
                cand = c;
                wk.j += c - pos;
                pos = c;
                settle(wk);
                ext = rec_of(wk.j, wk.base);
                strengths(wk.j, wk.base, same, s_own, s_ext);
            }
            double t, f[D];
            calc_err |= pair(pv, ext + q * D, ext + (q + 1) * D, same, s_own, f, t);
#pragma unroll
            for (int k = 0; k < D; k++) fo_sum[k] = fo_sum[k] + f[k];
        }
        if (cand != 0xffffffffu) {
#pragma unroll
            for (int k = 0; k < D; k++) acc[k] = acc[k] + fo_sum[k] * 0.5;
        }
    }

    // Ext part, calc_j(j, i) row v, for same-voxel candidates (every candidate in
    // the physical weighting): for each vertex u of j, the force on the nearest
    // point of i's segments (v - 1, v) (weight t) and (v, v + 1) (weight 1 - t).
    // Bit (c * 2 nv + 2 u + side) per chunk of candidates.
    const uint32_t per_e = 64u / (2u * nv) < 32u ? 64u / (2u * nv) : 32u;
    RodWalk ew;
    if (physical) {
        live = POT != CRM_POTENTIAL_NONE && first(ew);
    } else {
        ew.ra = ew.rb = 0;
        ew.j = cb;
        ew.je = ce;
        ew.base = row_base(idx[0], idx[1]);
        live = POT != CRM_POTENTIAL_NONE && ew.j < ew.je;
    }
    auto esettle = [&](RodWalk& wk) { return physical ? settle(wk) : wk.j < wk.je; };
    while (live) {
        const RodWalk c0 = ew;
        uint64_t me = 0;
        for (uint32_t c = 0; c < per_e; c++) {
            const uint32_t j = ew.j;
            if (j != s) {
                const double4 bj = bnd_of(j, ew.base);
                const double cj[3] = {bj.x, bj.y, bj.z};
                if (!out_of_reach(ci, cj, ri + (float)bj.w * (1.0f + 1e-6f))) {
                    const double* ext = rec_of(j, ew.base);
                    for (uint32_t u = 0; u < nv; u++) {
                        const double* pu = ext + u * D;
                        if (h0 >= 0.0f && !out_of_reach(pu, m0, h0)) me |= 1ull << (c * 2 * nv + 2 * u);
                        if (h1 >= 0.0f && !out_of_reach(pu, m1, h1)) me |= 1ull << (c * 2 * nv + 2 * u + 1);
                    }
                }
            }
            ++ew.j;
            live = esettle(ew);
            if (!live) break;
        }
        RodWalk wk = c0;
        uint32_t pos = 0, cand = 0xffffffffu;
        const double* ext = nullptr;
        bool same = false;
        double s_own = 0.0, s_ext = 0.0;
        double fe_sum[D];
#pragma unroll
        for (int k = 0; k < D; k++) fe_sum[k] = 0.0;
        while (me) {
            const uint32_t bit = (uint32_t)__builtin_ctzll(me);
            me &= me - 1;
            const uint32_t c = bit / (2 * nv), rem = bit - c * 2 * nv, u = rem >> 1, side = rem & 1u;
            if (c != cand) {
                if (cand != 0xffffffffu) {
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        acc[k] = acc[k] + fe_sum[k] * 0.5;
                        fe_sum[k] = 0.0;
                    }
                }
                cand = c;
                wk.j += c - pos;
                pos = c;
                if (physical) settle(wk);
                ext = rec_of(wk.j, wk.base);
                strengths(wk.j, wk.base, same, s_own, s_ext);
            }
            // side 0: segment (v - 1, v), v is its second end (weight t)
            // side 1: segment (v, v + 1), v is its first end (weight 1 - t)
            const int q = (int)v - 1 + (int)side;
            const double* pu = ext + u * D;
            double t, f[D];
            calc_err |= pair(pu, own + q * D, own + (q + 1) * D, same, s_ext, f, t);
            const double wgt = side == 0 ? t : 1.0 - t;
#pragma unroll
            for (int k = 0; k < D; k++) fe_sum[k] = fe_sum[k] + (-f[k]) * wgt;
        }
        if (cand != 0xffffffffu) {
#pragma unroll
            for (int k = 0; k < D; k++) acc[k] = acc[k] + fe_sum[k] * 0.5;
        }
    }
    if (calc_err) raise_error(A.ctrl, CRM_CALC_ERROR, A.ctrl->step, S.id[s]);
    // domain force after the pair forces (run_sim.rs:277-279, update_mechanics.rs:244-274)
    if (P.dforce == CRM_DOMAIN_FORCE_SPRING) {
#pragma unroll
        for (int k = 0; k < D; k++) acc[k] = acc[k] + (-P.dk[k]) * (pv[k] - P.dc[k]);
    } else if (D == 3 && P.dforce == CRM_DOMAIN_FORCE_ROD_GEL) {
        // CartesianSubDomainRods::calculate_custom_force (bacterial_rods.rs:447-477)
        double f[D], vv[D];
        bool near = false;
#pragma unroll
        for (int k = 0; k < D; k++) {
            f[k] = 0.0;
            vv[k] = S.vel[(uint64_t)(v * D + k) * A.N + s];
            near = near || fabs(pv[k] - P.min[k]) <= P.fric_d || fabs(pv[k] - P.max[k]) <= P.fric_d;
        }
        f[D - 1] = -P.gel;
        const double q = sqrt(sum_squares_unfused<D>(vv));
        if (q != 0.0 && near) {
#pragma unroll
            for (int k = 0; k < D; k++) f[k] = f[k] + ((-(vv[k] / q)) * P.gel) * P.fric;
        }
#pragma unroll
        for (int k = 0; k < D; k++) acc[k] = acc[k] + f[k];
    }
#pragma unroll
    for (int k = 0; k < D; k++) A.force[(uint64_t)(v * D + k) * A.N + s] = acc[k];

    // Neighbour sensing as in the rods example (bacterial_rods/src/main.rs:48-78):
    // vertex v counts the vertices u of every other rod of the stencil with
    // norm(v - u) < radius, i.e. squared distance below sense_t; rods whose
    // bounding ball is out of the radius' reach are skipped.
    if (A.sense) {
        const double tsq = A.sense_t;
        float sreach = (float)sqrt(tsq) * (1.0f + 1e-6f) + 1e-5f;
        if (!(sreach >= 0.0f)) sreach = __builtin_inff();
        uint32_t n = 0;
        RodWalk sw;
        bool more = first(sw);
        while (more) {
            const uint32_t j = sw.j;
            if (j != s) {
                const double4 bj = bnd_of(j, sw.base);
                const float lim = (float)bj.w * (1.0f + 1e-6f) + sreach;
                float d2 = 0.0f;
                const double cj[3] = {bj.x, bj.y, bj.z};
#pragma unroll
                for (int k = 0; k < D; k++) {
                    const float d = (float)(pv[k] - cj[k]);
                    d2 = fmaf(d, d, d2);
                }
                if (!(d2 > lim * lim)) {
                    const double* ext = rec_of(j, sw.base);
                    for (uint32_t u = 0; u < nv; u++) {
                        double z[D];
#pragma unroll
                        for (int k = 0; k < D; k++) z[k] = pv[k] - ext[u * D + k];
                        n += sum_squares_unfused<D>(z) < tsq ? 1u : 0u;
                    }
                }
            }
            ++sw.j;
            more = settle(sw);
        }
        A.sense[s * nv + v] = n;
    }
    };
    if (staged) work(std::true_type{});
    else work(std::false_type{});
}

// Normal k of a rod's K normals for (agent, step): Philox block k / 2, element k % 2.
__device__ __forceinline__ double rod_normal(const StepArgs& sa, uint32_t id, uint64_t t, uint32_t k) {
    if (sa.noise_mode == CRM_NOISE_BUFFER && sa.noise) return 0.0;  // caller supplies per-vertex values
    double z0, z1;
    gaussian_pair(sa.seed, id, t, k / 2, z0, z1);
    return (k & 1) ? z1 : z0;
}

// Bending window i of a rod (vertices i, i + 1, i + 2; bacterial_rods.rs:183-205):
// the force F on the middle vertex; the outer two get -F / 2 each.  pos(i, k)
// reads vertex i.  One lane per window computes it and shares it through LDS.
template <int D, typename Pos>
__device__ __forceinline__ void rod_bend_window(const InstParams& P, Pos pos, int i, double* F) {
    double c1[D], c2[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        c1[k] = pos(i, k) - pos(i + 1, k);
        c2[k] = pos(i + 1, k) - pos(i + 2, k);
    }
    double prod = 0.0, n1 = 0.0, n2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        prod = prod + c1[k] * c2[k];
        n1 = n1 + c1[k] * c1[k];
        n2 = n2 + c2[k] * c2[k];
    }
    n1 = sqrt(n1);
    n2 = sqrt(n2);
    double ang = 0.0;
    if (!(n1 == 0.0 || n2 == 0.0)) {  // nalgebra Matrix::angle
        double cang = prod / (n1 * n2);
        cang = cang < -1.0 ? -1.0 : (cang > 1.0 ? 1.0 : cang);
        ang = acos(cang);
    }
    double fd[D], nf = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        fd[k] = c1[k] - c2[k];
        nf = nf + fd[k] * fd[k];
    }
    nf = sqrt(nf);
    const double tn = tan(0.5 * ang);
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double fdir = nf != 0.0 ? fd[k] / nf : fd[k];
        F[k] = ((fdir * 2.0) * P.rigidity) * tn;
    }
}

// LDS of the bending windows of a workgroup's rods: lane t holds window v of its
// rod (t = lane of vertex v), read by the lanes of vertices v, v + 1, v + 2.
constexpr int kRodWinK = kBlock * 3;

// RodMechanics::calculate_increment row v (bacterial_rods.rs:157-213): springs,
// bending and damping, accumulated in the reference's order.  pos(i, k) and
// vel(i, k) read vertex i of the rod; win(i, k) component k of bending window i.
template <int D, typename Pos, typename Vel, typename Win>
__device__ __forceinline__ void rod_row_increment(const InstParams& P, double spring_length, uint32_t nv, uint32_t v,
                                                  Pos pos, Vel vel, Win win, const double* force, double* dx,
                                                  double* dv) {
    double total[D];
#pragma unroll
    for (int k = 0; k < D; k++) total[k] = force[k];
    // dist_internal[i] = pos[i] - pos[i + 1]
    auto dist = [&](int i, double* d) {
#pragma unroll
        for (int k = 0; k < D; k++) d[k] = pos(i, k) - pos(i + 1, k);
    };
    auto spring = [&](int i, double* fi) {  // force_internal of spring i (zero if degenerate)
        double d[D], n2 = 0.0;
        dist(i, d);
#pragma unroll
        for (int k = 0; k < D; k++) n2 = n2 + d[k] * d[k];
        const double nd = sqrt(n2);
        if (nd == 0.0) return false;
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] = ((-(d[k] / nd)) * P.spring_tension) * (nd - spring_length);
        return true;
    };
    // springs i = v - 1 (v is its second vertex) then i = v (v is its first vertex)
    if (v >= 1) {
        double fi[D];
        if (spring((int)v - 1, fi)) {
#pragma unroll
            for (int k = 0; k < D; k++) total[k] = total[k] + (-fi[k]) * 0.5;
        }
    }
    if (v + 1 < nv) {
        double fi[D];
        if (spring((int)v, fi)) {
#pragma unroll
            for (int k = 0; k < D; k++) total[k] = total[k] + fi[k] * 0.5;
        }
    }
    // bending windows i = v - 2, v - 1, v (v is vertex i + 2, i + 1, i)
    for (int i = (int)v - 2; i <= (int)v; i++) {
        if (i < 0 || i + 2 >= (int)nv) continue;
        const int role = (int)v - i;  // 0: v = i, 1: v = i + 1, 2: v = i + 2
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double F = win(i, k);
            total[k] = total[k] + (role == 1 ? F : (-F) * 0.5);
        }
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        dv[k] = total[k] - vel(v, k) * P.damping;
        dx[k] = vel(v, k);
    }
}

// Solver step of one vertex row (same formulas as point_step).
template <int D, bool WITH_VEL, typename Hist>
__device__ __forceinline__ void rod_row_solve(const StepArgs& sa, const Ring& rg, const double* x, const double* vv,
                                              const double* dx, const double* dv, const double* dxr, Hist hist,
                                              double* nx, double* nvel) {
    if (rg.cap == 0) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            nx[k] = (dx[k] * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) nvel[k] = (dv[k] * sa.dt + vv[k]) * 1.0 + 0.0 * sa.dt;
        }
    } else if (rg.n_read == 2) {
        const double f0 = 23.0 / 12.0, f1 = (-16.0) / 12.0, f2 = 5.0 / 12.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double h0 = rg.it[0] == rg.wslot ? dx[k] : hist(0, k, false);
            const double h1 = rg.it[1] == rg.wslot ? dx[k] : hist(1, k, false);
            nx[k] = ((dx[k] * f0 + (h0 * f1 + h1 * f2)) * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) {
                const double g0 = rg.it[0] == rg.wslot ? dv[k] : hist(0, k, true);
                const double g1 = rg.it[1] == rg.wslot ? dv[k] : hist(1, k, true);
                nvel[k] = ((dv[k] * f0 + (g0 * f1 + g1 * f2)) * sa.dt + vv[k]) * 1.0 + 0.0 * sa.dt;
            }
        }
    } else {
        const double g0c = 3.0 / 2.0, g1c = (-1.0) / 2.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double h0 = rg.it[0] == rg.wslot ? dx[k] : hist(0, k, false);
            nx[k] = ((dx[k] * g0c + h0 * g1c) * sa.dt + x[k]) * 1.0 + dxr[k] * sa.dt;
            if (WITH_VEL) {
                const double q0 = rg.it[0] == rg.wslot ? dv[k] : hist(0, k, true);
                nvel[k] = ((dv[k] * g0c + q0 * g1c) * sa.dt + vv[k]) * 1.0 + 0.0 * sa.dt;
            }
        }
    }
}

// Position noise of vertex row v: N(0, sqrt dt) * 2^-2 * D / dt (bacterial_rods.rs:215-240).
template <int D>
__device__ __forceinline__ void rod_row_noise(const InstParams& P, const StepArgs& sa, uint32_t id, uint64_t t,
                                              uint32_t nv, uint32_t v, double* dxr) {
#pragma unroll
    for (int k = 0; k < D; k++) dxr[k] = 0.0;
    bool have = false;
    double z[D];
    if (sa.noise_mode == CRM_NOISE_BUFFER && sa.noise) {
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = sa.noise[((uint64_t)id * nv + v) * D + k];
        have = true;
    } else if (sa.noise_mode == CRM_NOISE_PHILOX) {
#pragma unroll
        for (int k = 0; k < D; k++) z[k] = rod_normal(sa, id, t, v * D + k);
        have = true;
    }
    if (have) {
#pragma unroll
        for (int k = 0; k < D; k++) dxr[k] = (((0.0 + sa.sqrt_dt * z[k]) * 0.25) * P.diffusion_constant) / sa.dt;
    }
}

// Per-vertex reflection (bacterial_rods.rs:489-533); returns true if still outside.
template <int D, bool WITH_VEL>
__device__ __forceinline__ bool rod_row_reflect(const InstParams& P, double* nx, double* nvel) {
    bool out = false;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int ax = k + 3 - D;
        if (nx[k] < P.min[ax]) {
            nx[k] = P.min[ax] * 2.0 - nx[k];
            if (WITH_VEL) nvel[k] = fabs(nvel[k]);
        }
        if (nx[k] > P.max[ax]) {
            nx[k] = P.max[ax] * 2.0 - nx[k];
            if (WITH_VEL) nvel[k] = -fabs(nvel[k]);
        }
        out |= (nx[k] < P.min[ax]) || (nx[k] > P.max[ax]);
    }
    return out;
}

constexpr int kRodLdsK = kBlock * 3;  // one new position per lane

// Next centroid voxel of every rod.  The rod's lanes compute their new vertex
// positions (dx = velocity: the force does not enter the position update), the
// first lane sums them in vertex order and divides by nv (row_sum / nrows).
template <int D>
__global__ __launch_bounds__(kBlock) void k_key_rod(Arrays A, StateBuf S, RodArgs R, StepArgs sa) {
    if (halted(A.ctrl, A.ctrl->step)) return;
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
        const uint32_t id = S.id[s];
        const Ring rg = ring_at(t, sa.cap, sa.newest_first);
        double x[D], vv[D], dx[D], dv[D], dxr[D], nx[D], nvel[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            x[k] = S.pos[(uint64_t)(v * D + k) * N + s];
            vv[k] = S.vel[(uint64_t)(v * D + k) * N + s];
            dx[k] = vv[k];
            dv[k] = 0.0;
        }
        rod_row_noise<D>(P, sa, id, t, nv, v, dxr);
        auto hist = [&](int i, int k, bool) { return S.hdx[((uint64_t)rg.it[i] * nv * D + v * D + k) * N + s]; };
        rod_row_solve<D, false>(sa, rg, x, vv, dx, dv, dxr, hist, nx, nvel);
        outflag[threadIdx.x] = rod_row_reflect<D, false>(P, nx, nvel) ? 1u : 0u;
#pragma unroll
        for (int k = 0; k < D; k++) newpos[threadIdx.x * 3 + k] = nx[k];
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

// Per-rod fields that follow the rod through a sort or a re-layout (strides of
// the source and destination buffers may differ).
__device__ __forceinline__ void copy_rod_extras(const StateBuf& src, uint64_t o, const StateBuf& dst, uint64_t s,
                                                uint64_t n_src, uint64_t n_dst) {
    dst.spring[s] = src.spring[o];
    dst.birth[s] = src.birth[o];
    dst.growth[s] = src.growth[o];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        dst.ident[j * n_dst + s] = src.ident[j * n_src + o];
        dst.parent[j * n_dst + s] = src.parent[j * n_src + o];
    }
}

// Position of vertex row v after step t1 from its state at t1 (x, vel and the
// ring through hist), reflected (the position update of k_key_rod), into out[D].
template <int D, typename Hist>
__device__ __forceinline__ void rod_ahead(const InstParams& P, const StepArgs& sa, uint64_t t1, uint32_t id,
                                          uint32_t nv, uint32_t v, const double* x, const double* vel, Hist hist,
                                          double* out) {
    const Ring rg = ring_at(t1, sa.cap, sa.newest_first);
    double dx[D], dv[D], dxr[D], nx[D], unused[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        dx[k] = vel[k];
        dv[k] = 0.0;
    }
    rod_row_noise<D>(P, sa, id, t1, nv, v, dxr);
    rod_row_solve<D, false>(sa, rg, x, vel, dx, dv, dxr, hist, nx, unused);
    rod_row_reflect<D, false>(P, nx, unused);
#pragma unroll
    for (int k = 0; k < D; k++) out[k] = nx[k];
}

// Gather + integrate for rods: lane (s, v) computes vertex row v of the rod that
// lands in sorted slot s (coming from slot o = perm[s]).  With AHEAD it also bins
// the rod for the next step, as k_key_rod would after that step's forces: the
// position update needs neither the force nor noise other than the next step's
// (Philox, or none), so the new vertex positions, velocities and ring give the
// next centroid; the step's boundary error is raised here then.
template <int D, bool INTEGRATE, bool AHEAD>
__global__ __launch_bounds__(kBlock) void k_gather_rod(Arrays A, StateBuf src, StateBuf dst, RodArgs R,
                                                         StepArgs sa) {
    __shared__ double newpos[kRodLdsK];
    __shared__ double win[INTEGRATE ? kRodWinK : 1];
    __shared__ double aheadpos[AHEAD ? kRodLdsK : 1];
    if (A.ctrl->skip) return;
    const BlockWork w = A.work[blockIdx.x];
    const uint32_t nv = R.nv, K = nv * D;
    // rods fastest across the lanes: thread t takes vertex row t / rpb of the
    // block's rod t % rpb, so that a wave's loads and stores of one component row
    // cover consecutive slots.  lt = rod * nv + vertex addresses the LDS rows.  The
    // streams read or written once (forces, the new state and ring) bypass the
    // caches' retention (nontemporal; 200,000 2-D rods 125 -> 118 us); positions stay cached, the
    // neighbouring vertex rows read them again.
    const uint32_t rr = threadIdx.x % R.rpb, v = threadIdx.x / R.rpb;
    const uint64_t s = w.first + rr;
    const bool active = v < nv && rr < w.count;
    const uint32_t lt = rr * nv + v;
    const InstParams& P = A.ip[w.inst];
    const uint64_t N = A.N;
    double nx[D], nvel[D];
    if constexpr (INTEGRATE) {  // bending window v of the rod, shared with the next two lanes
        if (active && v + 2 < nv) {
            const uint32_t o = A.perm[s];
            auto pos = [&](int i, int k) { return src.pos[(uint64_t)(i * D + k) * N + o]; };
            double F[D];
            rod_bend_window<D>(P, pos, (int)v, F);
#pragma unroll
            for (int k = 0; k < D; k++) win[lt * 3 + k] = F[k];
        }
        __syncthreads();
    }
    if (active) {
        const uint32_t o = A.perm[s];
        const uint32_t id = src.id[o];
        if (INTEGRATE) {
            const uint64_t t = A.ctrl->step - 1;
            const Ring rg = ring_at(t, sa.cap, sa.newest_first);
            auto pos = [&](int i, int k) { return src.pos[(uint64_t)(i * D + k) * N + o]; };
            auto vel = [&](int i, int k) { return src.vel[(uint64_t)(i * D + k) * N + o]; };
            double x[D], vv[D], f[D], dx[D], dv[D], dxr[D];
#pragma unroll
            for (int k = 0; k < D; k++) {
                x[k] = pos((int)v, k);
                vv[k] = vel((int)v, k);
                f[k] = __builtin_nontemporal_load(&A.force[(uint64_t)(v * D + k) * N + o]);
            }
            auto wn = [&](int i, int k) { return win[(lt - v + (uint32_t)i) * 3 + k]; };
            rod_row_increment<D>(P, src.spring[o], nv, v, pos, vel, wn, f, dx, dv);
            rod_row_noise<D>(P, sa, id, t, nv, v, dxr);
            auto hist = [&](int i, int k, bool is_dv) {
                const double* h = is_dv ? src.hdv : src.hdx;
                return h[((uint64_t)rg.it[i] * K + v * D + k) * N + o];
            };
            rod_row_solve<D, true>(sa, rg, x, vv, dx, dv, dxr, hist, nx, nvel);
            const bool out = rod_row_reflect<D, true>(P, nx, nvel);
            if (AHEAD && out) raise_error(A.ctrl, CRM_BOUNDARY_ERROR, t, id);
            if constexpr (AHEAD) {
                // the ring after this step: slot rg.wslot holds dx, the others src's
                const Ring rg1 = ring_at(t + 1, sa.cap, sa.newest_first);
                auto hist1 = [&](int i, int k, bool) {
                    const uint32_t sl = rg1.it[i];
                    return sl == rg.wslot ? dx[k] : src.hdx[((uint64_t)sl * K + v * D + k) * N + o];
                };
                rod_ahead<D>(P, sa, t + 1, id, nv, v, nx, nvel, hist1, aheadpos + lt * 3);
            }
            if (rg.cap > 0) {
                const uint32_t n_hist = (uint32_t)min<uint64_t>(t + 1, rg.cap);
                for (uint32_t h = 0; h < n_hist; h++) {
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        const uint64_t e = ((uint64_t)h * K + v * D + k) * N;
                        __builtin_nontemporal_store(h == rg.wslot ? dx[k] : src.hdx[e + o], &dst.hdx[e + s]);
                        __builtin_nontemporal_store(h == rg.wslot ? dv[k] : src.hdv[e + o], &dst.hdv[e + s]);
                    }
                }
            }
        } else {
            // every ring slot: rods born at different steps hold different counts
            const uint32_t n_hist = sa.cap;
#pragma unroll
            for (int k = 0; k < D; k++) {
                nx[k] = src.pos[(uint64_t)(v * D + k) * N + o];
                nvel[k] = src.vel[(uint64_t)(v * D + k) * N + o];
            }
            for (uint32_t h = 0; h < n_hist; h++) {
#pragma unroll
                for (int k = 0; k < D; k++) {
                    const uint64_t e = ((uint64_t)h * K + v * D + k) * N;
                    dst.hdx[e + s] = src.hdx[e + o];
                    dst.hdv[e + s] = src.hdv[e + o];
                }
            }
            if constexpr (AHEAD) {
                const uint64_t t1 = A.ctrl->step;
                const Ring rg1 = ring_at(t1, sa.cap, sa.newest_first);
                auto hist1 = [&](int i, int k, bool) {
                    return src.hdx[((uint64_t)rg1.it[i] * K + v * D + k) * N + o];
                };
                rod_ahead<D>(P, sa, t1, id, nv, v, nx, nvel, hist1, aheadpos + lt * 3);
            }
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            __builtin_nontemporal_store(nx[k], &dst.pos[(uint64_t)(v * D + k) * N + s]);
            __builtin_nontemporal_store(nvel[k], &dst.vel[(uint64_t)(v * D + k) * N + s]);
            newpos[lt * 3 + k] = nx[k];
        }
        if (v == 0) {
            dst.species[s] = src.species[o];
            dst.id[s] = id;
            dst.vox[s] = A.key[o];
            copy_rod_extras(src, o, dst, s, N, N);
        }
    }
    __syncthreads();
    // the force records rec[slot][K] of the block's rods, a contiguous range,
    // written element by element from LDS
    for (uint32_t e = threadIdx.x; e < w.count * K; e += kBlock) {
        const uint32_t r = e / K, c = e - r * K, vq = c / D;
        R.rec[w.first * K + e] = newpos[(r * nv + vq) * 3 + (c - vq * D)];
    }
    // the rod's vertex mean (every lane of the rod computes the same)
    auto mean = [&](double* c) {
        const uint32_t l0 = rr * nv;
        c[0] = c[1] = c[2] = 0.0;
        for (uint32_t q = 0; q < nv; q++) {
#pragma unroll
            for (int k = 0; k < D; k++) c[k] = c[k] + newpos[(l0 + q) * 3 + k];
        }
#pragma unroll
        for (int k = 0; k < D; k++) c[k] = c[k] / (double)nv;
    };
    // ball of segment (v, v + 1): midpoint relative to the mean, half length with
    // margin (the pass-1 reach tests of the force kernel)
    if (active && v + 1 < nv) {
        double c[3];
        mean(c);
        float o[3] = {0.0f, 0.0f, 0.0f}, h2 = 0.0f;
        bool zero = true;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double a = newpos[lt * 3 + k], b = newpos[(lt + 1) * 3 + k];
            o[k] = (float)((a + b) * 0.5 - c[k]);
            const float d = (float)(b - a);
            h2 = fmaf(d, d, h2);
            zero = zero && a == b;
        }
        const float h = zero ? __builtin_inff() : 0.5f * sqrtf(h2) * (1.0f + 1e-6f) + 1e-6f;
        R.seg[s * (nv - 1) + v] = make_float4(o[0], o[1], o[2], h);
    }
    // bounding ball of the rod (first lane): vertex mean and the largest distance to it
    if (active && v == 0) {
        double c[3];
        mean(c);
        double r2 = 0.0;
        for (uint32_t q = 0; q < nv; q++) {
            double d2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) {
                const double d = newpos[(lt + q) * 3 + k] - c[k];
                d2 = d2 + d * d;
            }
            r2 = fmax(r2, d2);
        }
        // a zero-length segment gets an infinite radius: such a rod is never
        // skipped, so its undefined nearest points reach the sums as in the
        // reference
        bool degenerate = false;
        for (uint32_t q = 0; q + 1 < nv; q++) {
            bool same = true;
#pragma unroll
            for (int k = 0; k < D; k++)
                same = same && newpos[(lt + q) * 3 + k] == newpos[(lt + q + 1) * 3 + k];
            degenerate = degenerate || same;
        }
        // radius rounded up by a relative 1e-12 so that it covers every vertex
        const double radius = degenerate ? __builtin_inf() : sqrt(r2) * (1.0 + 1e-12);
        R.bnd[s] = make_double4(c[0], c[1], D == 3 ? c[2] : 0.0, radius);
        if constexpr (AHEAD) {
            // next voxel: the centroid after the next step (row_sum / nrows), as k_key_rod
            double c1[D];
#pragma unroll
            for (int k = 0; k < D; k++) c1[k] = 0.0;
            for (uint32_t q = 0; q < nv; q++) {
#pragma unroll
                for (int k = 0; k < D; k++) c1[k] = c1[k] + aheadpos[(lt + q) * 3 + k];
            }
            uint32_t idx[3] = {0, 0, 0};
            bool ok = true;
#pragma unroll
            for (int k = 0; k < D; k++) {
                const int ax = k + 3 - D;
                ok = ok && axis_index(c1[k] / (double)nv, P.min[ax], P.dx[ax], P.n[ax], idx[ax]);
            }
            uint32_t key = dst.vox[s];
            if (ok) key = P.voxel_offset + (idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2];
            else atomicAdd((unsigned long long*)&A.ctrl->out_of_grid, 1ull);
            A.key_next[s] = key;
        }
    }
}

}  // namespace crm
