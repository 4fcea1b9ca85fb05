// crmech.hip -- host runtime and C ABI of the gfx950 mechanics backend.
//
// One context = one HIP stream, one set of device buffers, one or more
// independent instances (CartesianCuboid domains) advanced together.
// The drop-in boundary is include/crmech.h; see there for the reference
// interfaces each entry point replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <filesystem>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "crmech.h"
#include "kernels_rod.hpp"
#include "kernels_snapshot.hpp"
#include "kernels_cycle.hpp"
#include "kernels_remove.hpp"

#include <hipcub/hipcub.hpp>

namespace crm {

#define CRM_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ctx->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr;     \
            return CRM_DEVICE_ERROR;                                                         \
        }                                                                                    \
    } while (0)

struct HostGrid {
    double min[3], max[3], dx[3];
    uint64_t n[3];  // 3-D layout, axis 0 extent 1 for 2-D
    uint64_t V;
};

// num-traits to_usize on f64 (truncation, -1 < x < 2^64, NaN fails)
static bool to_usize(double x, uint64_t& out) {
    if (std::isnan(x) || !(x > -1.0) || !(x < 18446744073709551616.0)) return false;
    double t = std::trunc(x);
    out = t <= 0.0 ? 0 : (uint64_t)t;
    return true;
}

// CartesianCuboid::from_boundaries_and_interaction_range / _n_voxels
// (cartesian_cuboid_n.rs:149-162,183-256).
static crm_status host_grid(const crm_domain& d, uint32_t D, HostGrid& g, std::string& err) {
    if (D < 2 || D > 3) {
        err = "dim must be 2 or 3";
        return CRM_INVALID_ARGUMENT;
    }
    for (uint32_t i = 0; i < D; i++)
        if (d.min[i] >= d.max[i]) {
            err = "BoundaryError: Min must be smaller than Max for domain boundaries";
            return CRM_BOUNDARY_ERROR;
        }
    for (int a = 0; a < 3; a++) {
        g.min[a] = 0.0;
        g.max[a] = 1.0;
        g.dx[a] = 1.0;
        g.n[a] = 1;
    }
    for (uint32_t i = 0; i < D; i++) {
        const int a = (int)(i + 3 - D);
        g.min[a] = d.min[i];
        g.max[a] = d.max[i];
        if (d.n_voxels[0] == 0) {
            double n = std::floor((d.max[i] - d.min[i]) / d.interaction_range);
            uint64_t m;
            if (!to_usize(n, m)) {
                err = "BoundaryError: Cannot convert float to usize (conversion error during domain setup)";
                return CRM_BOUNDARY_ERROR;
            }
            g.n[a] = std::max<uint64_t>(m, 1);
            g.dx[a] = (d.max[i] - d.min[i]) / n;
        } else {
            if (d.n_voxels[i] == 0) {
                err = "BoundaryError: n_voxels must be positive";
                return CRM_BOUNDARY_ERROR;
            }
            g.n[a] = d.n_voxels[i];
            g.dx[a] = (d.max[i] - d.min[i]) / (double)d.n_voxels[i];
        }
    }
    g.V = g.n[0] * g.n[1] * g.n[2];
    return CRM_OK;
}

// get_decomp_res (cartesian_cuboid_n.rs:19-59): V = avg * n + (avg - 1) * m with
// n + m = k chunks, searched the reference's way (avg from ceil(V / k) upwards,
// trading full chunks for short ones).  False when no split is found.
static bool decomp_res(uint64_t V, uint64_t k, uint64_t& n_out, uint64_t& avg_out) {
    int64_t avg = (int64_t)std::ceil((double)V / (double)k);
    int64_t n = (int64_t)k, m = 0;
    for (uint64_t it = 0; it < k; it++) {
        const int64_t rest = (int64_t)V - avg * n - (avg - 1) * m;
        if (rest == 0) {
            n_out = (uint64_t)n;
            avg_out = (uint64_t)avg;
            return true;
        }
        if (rest > 0) {
            if (n == (int64_t)k) {
                avg += 1;
                m = 0;
            }
        } else {
            n -= 1;
            m += 1;
        }
    }
    return false;
}

// Interactions whose parameters depend on the agents' species.
static bool species_kind(uint32_t kind) {
    return kind == CRM_POTENTIAL_MORSE_SPECIES || kind == CRM_POTENTIAL_CELL_SORTING ||
           kind == CRM_POTENTIAL_CELL_SORTING_LANGEVIN;
}

static int exact_int_exponent(double e) {
    if (e >= 0.0 && e <= 64.0 && std::floor(e) == e) return (int)e;
    return -1;
}

template <typename T>
static void fill_pot(PotParams<T>& q, const crm_potential& p) {
    const double R2 = p.radius + p.radius;
    q.radius = (T)p.radius;
    q.two_radius = (T)R2;
    q.stiffness = (T)p.potential_stiffness;
    q.cutoff = (T)p.cutoff;
    q.strength = (T)p.strength;
    q.bound = (T)p.bound;
    q.en = (T)p.en;
    q.em = (T)p.em;
    // MiePotential: sigma = (em/en)^(1/(en-em)) * (R_own + R_ext) and
    // C = en/(en-em) * (en/em)^(em/(en-em)) (interaction.rs:363-367,390-392)
    q.mie_sigma = (T)(std::pow(p.em / p.en, 1.0 / (p.en - p.em)) * R2);
    q.mie_constant = (T)(p.en / (p.en - p.em) * std::pow(p.en / p.em, p.em / (p.en - p.em)));
    q.mie_en_s = q.en / q.mie_sigma;
    q.mie_em_s = q.em / q.mie_sigma;
    q.epsilon = (T)p.epsilon;
    q.sigma = (T)p.sigma;
    q.rel_range = (T)p.relative_interaction_range;
    q.sign = (T)(p.sign < 0 ? -1.0 : 1.0);
    for (int i = 0; i < CRM_MAX_SPECIES * CRM_MAX_SPECIES; i++) q.species_matrix[i] = (T)p.species_matrix[i];
    q.ien1 = exact_int_exponent(p.en + 1.0);
    q.iem1 = exact_int_exponent(p.em + 1.0);
    q.pad0 = q.pad1 = 0;
}

enum { kNumKernelClasses = CRM_KERNEL_COUNT };

// ---- device scratch of crm_remove_agents_noalloc ---------------------------
// One static array per GPU, sized for kNoAllocMaxAgents agents in at most
// kNoAllocMaxInstances instances: flags [n] bytes, four [n + 1] uint32 arrays
// (survivor flags by slot and by agent index, their exclusive scans), removed
// agents per instance, and the scan's temporary storage, at 256-byte aligned
// offsets.  Host code uses the per-device address from hipGetSymbolAddress,
// never the symbol itself.
constexpr uint64_t kNoAllocMaxAgents = 1ull << 22;
constexpr uint32_t kNoAllocMaxInstances = 1u << 16;
constexpr size_t kNoAllocScanReserve = (size_t)1 << 20;
// Steps per long graph replay (even, so that a replay keeps the buffer parity).
constexpr uint32_t kLongGraphSteps = 16;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocScratchBytes = align256(kNoAllocMaxAgents) + 4 * align256(4 * (kNoAllocMaxAgents + 1)) +
                                        align256(4 * (size_t)kNoAllocMaxInstances) + kNoAllocScanReserve;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

}  // namespace crm

using namespace crm;

struct crm_ctx {
    crm_config cfg;
    uint32_t D = 3, nv = 1, K = 3;
    uint32_t n_inst = 0;
    std::vector<crm_instance> inst;
    std::vector<HostGrid> grids;
    std::vector<InstParams> hparams;
    uint64_t Vt = 0;
    uint32_t pot_kind = 0, mech_kind = 0;
    bool symmetric = true;
    std::string err;

    int device = 0;
    hipStream_t stream = nullptr;

    // device memory
    InstParams* d_ip = nullptr;
    BlockWork* d_work = nullptr;
    uint32_t n_work = 0;
    ForceTile* d_tiles = nullptr;
    uint32_t n_force_tiles = 0;
    size_t tiles_cap = 0;
    uint32_t policy = 0;                  // tiling policy (index into kTilePolicies)
    uint32_t lanes_log2 = 0;              // lanes per own agent in the force kernel (policy's)
    std::vector<uint32_t> tile_rows, tile_cols, tile_count;  // per instance
    Control* d_ctrl = nullptr;
    uint32_t* d_start = nullptr;
    uint32_t* d_count = nullptr;
    uint32_t* d_start_local = nullptr;  // [Vt] voxel offsets inside their scan tile
    uint32_t* d_bsum = nullptr;         // [n_tiles] scan tile sums, then offsets
    uint32_t n_tiles = 0;
    uint32_t* d_key = nullptr;
    uint32_t* d_rank = nullptr;
    uint32_t* d_perm = nullptr;
    double* d_force = nullptr;
    double4* d_rec64 = nullptr;
    int4* d_recq = nullptr;
    StateBuf buf[2] = {};
    int cur = 0;
    double* d_noise = nullptr;
    size_t noise_cap = 0;
    double* d_recrod = nullptr;  // rods: AoS vertex positions [N][K]
    double4* d_bnd = nullptr;    // rods: bounding balls [N]
    float4* d_seg = nullptr;     // rods: segment balls [N][nv - 1]
    uint32_t rods_per_block = 0;
    // neighbour sensing: counts by slot ([N][nv] for rods) of the last force
    // evaluation, which read state buffer sense_parity (-1: none since enabled)
    bool sense_on = false;
    double sense_t = 0.0;
    uint32_t* d_sense = nullptr;
    int sense_parity = -1;
    // rods example cycle (crm_set_rod_cycle): capacity, per-slot and per-voxel scratch
    bool cycle_on = false;
    crm_rod_cycle cyc{};
    uint64_t cap_agents = 0;        // rods allocated in every N-sized array
    uint32_t* d_divflag = nullptr;  // [cap] 1: divides this step
    uint32_t* d_drank = nullptr;    // [cap] rank of a dividing rod among its voxel's
    uint64_t* d_dctr = nullptr;     // [cap] identifier counter of its daughter
    uint32_t* d_ndiv = nullptr;     // [Vt + 1] daughters per voxel, then their exclusive scan
    uint32_t* d_dbefore = nullptr;  // [Vt + 1]
    uint64_t* d_vcount = nullptr;   // [Vt] Voxel::id_counter
    uint32_t* d_inst_div = nullptr; // [n_inst] dividing rods per instance
    uint8_t* d_rm = nullptr;        // agent removal scratch (grow-only), rm_cap bytes
    size_t rm_cap = 0;
    uint8_t* d_noalloc = nullptr;   // this device's copy of `scratch` (crm_remove_agents_noalloc)
    std::vector<uint32_t> inst_host;  // [n_inst] host copy of per-instance counters
    bool graphs_stale = false;      // captured graphs no longer match (destroyed by the next crm_step)
    std::vector<BlockWork> work_host;  // the work table last uploaded
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    // snapshot records (agent order) and species, allocated on the first save
    double* d_snap = nullptr;
    int32_t* d_snap_sp = nullptr;
    double* d_snap_rod = nullptr;     // [N][2] rods: spring length, growth rate
    uint64_t* d_snap_id = nullptr;    // [N][6] rods: identifier, parent
    uint32_t* d_snap_st = nullptr;    // [N] rods: history entries held
    size_t snap_cap = 0;

    uint64_t N = 0;
    std::vector<uint64_t> counts;
    // point agents: CellIdentifier and parent by agent index, (kind, a, b) each (rods
    // keep theirs on the device); renumbered with the agent indices on removal
    std::vector<uint64_t> pt_ident, pt_parent;
    std::vector<uint64_t> pt_idc;  // points: Voxel::id_counter of every voxel (divisions number from it)
    bool pt_births = false;        // points: buf[].birth allocated (agents inserted by crm_insert_agents)
    bool have_agents = false;
    bool halted = false;
    uint64_t step = 0;

    // graphs for the two buffer parities: [parity] replays two steps, [2 + parity]
    // kLongGraphSteps steps
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};

    // timing: an event pair around each launch (class t_cls[i]), recorded without
    // waiting, so that the timed kernels run back to back as in a normal step;
    // the pairs are read when the pool is full and when the timings are read
    static constexpr int kTimingPairs = 64;
    hipEvent_t ev_a[kTimingPairs] = {}, ev_b[kTimingPairs] = {};
    int t_cls[kTimingPairs] = {};
    int t_used = 0;
    double t_ms[CRM_KERNEL_COUNT] = {};
    uint64_t t_n[CRM_KERNEL_COUNT] = {};
};

namespace {

// Message of the last failed crm_create on this thread (crm_last_error(NULL, ...)).
thread_local std::string g_create_error;

// Address of `scratch` on each device, looked up once per device (with that
// device current) by the first crm_create there.
std::mutex g_scratch_mutex;
std::map<int, uint8_t*> g_scratch_ptr;

hipError_t scratch_address(int device, uint8_t*& out) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = out = static_cast<uint8_t*>(p);
    return hipSuccess;
}

void free_device(crm_ctx* ctx) {
    // the graph executables first: their kernel nodes hold pointers into the buffers
    // below, so they must not outlive them
    for (auto& g : ctx->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    ctx->graphs_stale = false;
    void* ptrs[] = {ctx->d_work, ctx->d_tiles, ctx->d_start, ctx->d_count, ctx->d_start_local, ctx->d_bsum,
                    ctx->d_key, ctx->d_rank, ctx->d_perm, ctx->d_force, ctx->d_rec64, ctx->d_recq, ctx->d_noise, ctx->d_recrod, ctx->d_bnd,
                    ctx->d_seg, ctx->d_sense, ctx->d_snap, ctx->d_snap_sp, ctx->d_snap_rod, ctx->d_snap_id, ctx->d_snap_st, ctx->d_divflag, ctx->d_drank, ctx->d_dctr,
                    ctx->d_ndiv, ctx->d_dbefore, ctx->d_vcount, ctx->d_scan_tmp, ctx->d_inst_div, ctx->d_rm};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& b : ctx->buf) {
        void* q[] = {b.pos, b.vel, b.hdx, b.hdv, b.species, b.id, b.vox, b.spring, b.birth, b.growth, b.ident, b.parent};
        for (void* p : q)
            if (p) (void)hipFree(p);
        b = StateBuf{};
    }
    ctx->d_work = nullptr;
    ctx->d_tiles = nullptr;
    ctx->tiles_cap = 0;
    ctx->d_start = ctx->d_count = ctx->d_start_local = ctx->d_bsum = ctx->d_key = ctx->d_rank = ctx->d_perm = nullptr;
    ctx->d_force = nullptr;
    ctx->d_rec64 = nullptr;
    ctx->d_recq = nullptr;
    ctx->d_noise = nullptr;
    ctx->d_recrod = nullptr;
    ctx->d_bnd = nullptr;
    ctx->d_seg = nullptr;
    ctx->d_sense = nullptr;
    ctx->sense_parity = -1;
    ctx->d_snap = nullptr;
    ctx->d_snap_sp = nullptr;
    ctx->d_snap_rod = nullptr;
    ctx->d_snap_id = nullptr;
    ctx->d_snap_st = nullptr;
    ctx->snap_cap = 0;
    ctx->d_divflag = ctx->d_drank = nullptr;
    ctx->d_dctr = nullptr;
    ctx->d_ndiv = ctx->d_dbefore = nullptr;
    ctx->d_vcount = nullptr;
    ctx->d_inst_div = nullptr;
    ctx->d_rm = nullptr;
    ctx->rm_cap = 0;
    ctx->d_scan_tmp = nullptr;
    ctx->scan_tmp_bytes = 0;
    ctx->cap_agents = 0;
    ctx->noise_cap = 0;
}

// d_count holds the arrivals per voxel, then (from this offset, 256-byte aligned)
// the stayers per voxel.
uint64_t stay_offset(const crm_ctx* ctx) { return (ctx->Vt + 63) & ~(uint64_t)63; }

bool force_f32(const crm_ctx* ctx);

Arrays arrays(crm_ctx* ctx) {
    Arrays A;
    A.ip = ctx->d_ip;
    A.work = ctx->d_work;
    A.ctrl = ctx->d_ctrl;
    A.start = ctx->d_start;
    A.start_local = ctx->d_start_local;
    A.tile_off = ctx->d_bsum;
    A.count = ctx->d_count;
    A.stay = ctx->d_count + stay_offset(ctx);
    // two key buffers: the current step's (read by the sort) and the next step's
    // (binned ahead by the gather), by state-buffer parity
    A.key = ctx->d_key + (size_t)ctx->cur * ctx->N;
    A.rank = ctx->d_rank;
    A.key_next = ctx->d_key + (size_t)(1 - ctx->cur) * ctx->N;
    A.perm = ctx->d_perm;
    A.force = ctx->d_force;
    A.forcef = force_f32(ctx) ? reinterpret_cast<float*>(ctx->d_force) : nullptr;
    A.rec64 = ctx->d_rec64;
    A.recq = ctx->d_recq;
    A.sense = ctx->sense_on ? ctx->d_sense : nullptr;
    A.sense_t = ctx->sense_t;
    A.N = ctx->N;
    A.Vt = (uint32_t)ctx->Vt;
    return A;
}

bool is_rod(const crm_ctx* ctx) { return ctx->mech_kind == CRM_MECHANICS_ROD; }

// Point agents in mixed precision keep their forces in float (the pair sums are
// float already); the buffer keeps its double size.
bool force_f32(const crm_ctx* ctx) { return !is_rod(ctx) && ctx->cfg.precision == CRM_PRECISION_MIXED; }

RodArgs rod_args(crm_ctx* ctx) {
    RodArgs R;
    R.rec = ctx->d_recrod;
    R.bnd = ctx->d_bnd;
    R.seg = ctx->d_seg;
    R.nv = ctx->nv;
    R.rpb = ctx->rods_per_block;
    return R;
}

bool timing_on(crm_ctx* ctx) { return (ctx->cfg.flags & CRM_FLAG_TIMING) != 0; }

// Waits for the recorded event pairs and adds their times to their classes.
crm_status t_flush(crm_ctx* ctx) {
    if (ctx->t_used == 0) return CRM_OK;
    const int n = ctx->t_used;
    ctx->t_used = 0;
    CRM_HIP(hipEventSynchronize(ctx->ev_b[n - 1]));
    for (int i = 0; i < n; i++) {
        float ms = 0.f;
        CRM_HIP(hipEventElapsedTime(&ms, ctx->ev_a[i], ctx->ev_b[i]));
        ctx->t_ms[ctx->t_cls[i]] += ms;
        ctx->t_n[ctx->t_cls[i]] += 1;
    }
    return CRM_OK;
}

crm_status t_begin(crm_ctx* ctx) {
    if (timing_on(ctx)) {
        if (ctx->t_used == crm_ctx::kTimingPairs) {
            const crm_status st = t_flush(ctx);
            if (st != CRM_OK) return st;
        }
        CRM_HIP(hipEventRecord(ctx->ev_a[ctx->t_used], ctx->stream));
    }
    return CRM_OK;
}

crm_status t_end(crm_ctx* ctx, int cls) {
    CRM_HIP(hipGetLastError());
    if (timing_on(ctx)) {
        CRM_HIP(hipEventRecord(ctx->ev_b[ctx->t_used], ctx->stream));
        ctx->t_cls[ctx->t_used++] = cls;
    }
    return CRM_OK;
}

#define CRM_TRY(expr)                  \
    do {                               \
        crm_status s_ = (expr);        \
        if (s_ != CRM_OK) return s_;   \
    } while (0)

template <int D, int POT, bool MULTI>
crm_status launch_force_dpm(crm_ctx* ctx, StateBuf S, uint32_t rank) {
    Arrays A = arrays(ctx);
    const uint32_t phys = ctx->cfg.weighting == CRM_WEIGHT_PHYSICAL;
    const uint32_t sym = ctx->symmetric;
    const uint32_t no_lds = (ctx->cfg.flags & CRM_FLAG_NO_LDS) ? 1u : 0u;
    const dim3 g(ctx->n_force_tiles), b(kBlock);
    const uint32_t lg = ctx->lanes_log2;
    // sensing has kernels of its own, so the force-only kernels carry none of it
    const bool mixed = ctx->cfg.precision == CRM_PRECISION_MIXED, sense = A.sense != nullptr;
    if (mixed && sense)
        hipLaunchKernelGGL((k_force_tiled<D, POT, true, true, MULTI>), g, b, 0, ctx->stream, A, S, ctx->d_tiles, phys,
                           sym, no_lds, rank, lg);
    else if (mixed)
        hipLaunchKernelGGL((k_force_tiled<D, POT, true, false, MULTI>), g, b, 0, ctx->stream, A, S, ctx->d_tiles, phys,
                           sym, no_lds, rank, lg);
    else if (sense)
        hipLaunchKernelGGL((k_force_tiled<D, POT, false, true, MULTI>), g, b, 0, ctx->stream, A, S, ctx->d_tiles, phys,
                           sym, no_lds, rank, lg);
    else
        hipLaunchKernelGGL((k_force_tiled<D, POT, false, false, MULTI>), g, b, 0, ctx->stream, A, S, ctx->d_tiles,
                           phys, sym, no_lds, rank, lg);
    return CRM_OK;
}

// rank != 0: the force kernel also ranks the step's re-binning (keys binned ahead).
template <int D, int POT>
crm_status launch_force_dp(crm_ctx* ctx, StateBuf S, uint32_t rank) {
    return ctx->lanes_log2 ? launch_force_dpm<D, POT, true>(ctx, S, rank) : launch_force_dpm<D, POT, false>(ctx, S, rank);
}

template <int D>
crm_status launch_force_d(crm_ctx* ctx, StateBuf S, uint32_t rank) {
    switch (ctx->pot_kind) {
        case CRM_POTENTIAL_MORSE: return launch_force_dp<D, CRM_POTENTIAL_MORSE>(ctx, S, rank);
        case CRM_POTENTIAL_MIE: return launch_force_dp<D, CRM_POTENTIAL_MIE>(ctx, S, rank);
        case CRM_POTENTIAL_BOUND_LENNARD_JONES:
            return launch_force_dp<D, CRM_POTENTIAL_BOUND_LENNARD_JONES>(ctx, S, rank);
        case CRM_POTENTIAL_MORSE_SPECIES: return launch_force_dp<D, CRM_POTENTIAL_MORSE_SPECIES>(ctx, S, rank);
        case CRM_POTENTIAL_CELL_SORTING: return launch_force_dp<D, CRM_POTENTIAL_CELL_SORTING>(ctx, S, rank);
        case CRM_POTENTIAL_NONE: return launch_force_dp<D, CRM_POTENTIAL_NONE>(ctx, S, rank);
        case CRM_POTENTIAL_CELL_SORTING_LANGEVIN:
            return launch_force_dp<D, CRM_POTENTIAL_CELL_SORTING_LANGEVIN>(ctx, S, rank);
    }
    return CRM_INVALID_ARGUMENT;
}

template <int D, bool MIXED>
crm_status launch_force_rod_dm(crm_ctx* ctx, StateBuf S) {
    Arrays A = arrays(ctx);
    const RodArgs R = rod_args(ctx);
    const uint32_t phys = ctx->cfg.weighting == CRM_WEIGHT_PHYSICAL;
    const uint32_t no_lds = (ctx->cfg.flags & CRM_FLAG_NO_LDS) ? 1u : 0u;
    const dim3 g(ctx->n_work), b(kBlock);
    switch (ctx->pot_kind) {
        case CRM_POTENTIAL_MORSE:
            hipLaunchKernelGGL((k_force_rod<D, CRM_POTENTIAL_MORSE, MIXED>), g, b, 0, ctx->stream, A, S, R, phys,
                               no_lds);
            return CRM_OK;
        case CRM_POTENTIAL_MORSE_SPECIES:
            hipLaunchKernelGGL((k_force_rod<D, CRM_POTENTIAL_MORSE_SPECIES, MIXED>), g, b, 0, ctx->stream, A, S, R,
                               phys, no_lds);
            return CRM_OK;
        case CRM_POTENTIAL_MIE:
            hipLaunchKernelGGL((k_force_rod<D, CRM_POTENTIAL_MIE, MIXED>), g, b, 0, ctx->stream, A, S, R, phys,
                               no_lds);
            return CRM_OK;
        case CRM_POTENTIAL_BOUND_LENNARD_JONES:
            hipLaunchKernelGGL((k_force_rod<D, CRM_POTENTIAL_BOUND_LENNARD_JONES, MIXED>), g, b, 0, ctx->stream, A, S,
                               R, phys, no_lds);
            return CRM_OK;
        case CRM_POTENTIAL_NONE:
            hipLaunchKernelGGL((k_force_rod<D, CRM_POTENTIAL_NONE, MIXED>), g, b, 0, ctx->stream, A, S, R, phys,
                               no_lds);
            return CRM_OK;
    }
    return CRM_INVALID_ARGUMENT;
}

template <int D>
crm_status launch_force_rod_d(crm_ctx* ctx, StateBuf S) {
    return ctx->cfg.precision == CRM_PRECISION_MIXED ? launch_force_rod_dm<D, true>(ctx, S)
                                                     : launch_force_rod_dm<D, false>(ctx, S);
}

// rank: point agents binned ahead -- the force kernel also ranks the step's
// re-binning (only inside a step: a force evaluation on its own must leave the
// histogram alone).
crm_status launch_force(crm_ctx* ctx, bool rank = false) {
    CRM_TRY(t_begin(ctx));
    StateBuf S = ctx->buf[ctx->cur];
    ctx->sense_parity = ctx->cur;
    crm_status st;
    const uint32_t r = rank ? 1u : 0u;
    if (is_rod(ctx)) st = ctx->D == 2 ? launch_force_rod_d<2>(ctx, S) : launch_force_rod_d<3>(ctx, S);
    else st = ctx->D == 2 ? launch_force_d<2>(ctx, S, r) : launch_force_d<3>(ctx, S, r);
    if (st) return st;
    return t_end(ctx, CRM_KERNEL_FORCE);
}

StepArgs step_args(crm_ctx* ctx, const double* noise_step) {
    StepArgs sa;
    sa.noise = noise_step;
    sa.seed = ctx->cfg.seed;
    sa.dt = ctx->cfg.dt;
    sa.sqrt_dt = std::sqrt(ctx->cfg.dt);
    sa.noise_mode = ctx->cfg.noise_mode;
    sa.cap = ctx->cfg.solver_order;
    sa.newest_first = ctx->cfg.history_order == CRM_HISTORY_NEWEST_FIRST;
    sa.pad = 0;
    return sa;
}

template <int D>
crm_status launch_key_d(crm_ctx* ctx, StateBuf S, const StepArgs& sa) {
    Arrays A = arrays(ctx);
    switch (ctx->mech_kind) {
        case CRM_MECHANICS_NEWTON_DAMPED:
            hipLaunchKernelGGL((k_key_point<D, CRM_MECHANICS_NEWTON_DAMPED>), dim3(ctx->n_work), dim3(kBlock), 0,
                               ctx->stream, A, S, sa);
            return CRM_OK;
        case CRM_MECHANICS_LANGEVIN:
            hipLaunchKernelGGL((k_key_point<D, CRM_MECHANICS_LANGEVIN>), dim3(ctx->n_work), dim3(kBlock), 0,
                               ctx->stream, A, S, sa);
            return CRM_OK;
        case CRM_MECHANICS_BROWNIAN:
            hipLaunchKernelGGL((k_key_point<D, CRM_MECHANICS_BROWNIAN>), dim3(ctx->n_work), dim3(kBlock), 0,
                               ctx->stream, A, S, sa);
            return CRM_OK;
    }
    return CRM_INVALID_ARGUMENT;
}

// Ranks of the keys of the current order (stayers in order, arrivals atomic).
crm_status launch_rank(crm_ctx* ctx) {
    hipLaunchKernelGGL(k_rank_ahead, dim3((uint32_t)((ctx->Vt + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       arrays(ctx));
    CRM_HIP(hipGetLastError());
    return CRM_OK;
}

crm_status launch_key(crm_ctx* ctx, const StepArgs& sa) {
    CRM_TRY(t_begin(ctx));
    StateBuf S = ctx->buf[ctx->cur];
    crm_status st = CRM_OK;
    if (is_rod(ctx)) {
        Arrays A = arrays(ctx);
        const RodArgs R = rod_args(ctx);
        if (ctx->D == 2)
            hipLaunchKernelGGL((k_key_rod<2>), dim3(ctx->n_work), dim3(kBlock), 0, ctx->stream, A, S, R, sa);
        else
            hipLaunchKernelGGL((k_key_rod<3>), dim3(ctx->n_work), dim3(kBlock), 0, ctx->stream, A, S, R, sa);
    } else {
        st = ctx->D == 2 ? launch_key_d<2>(ctx, S, sa) : launch_key_d<3>(ctx, S, sa);
    }
    if (st) return st;
    CRM_TRY(launch_rank(ctx));
    return t_end(ctx, CRM_KERNEL_INTEGRATE);
}

template <int D, bool MIXED>
crm_status launch_gather_integrate_dm(crm_ctx* ctx, StateBuf src, StateBuf dst, const StepArgs& sa) {
    Arrays A = arrays(ctx);
    switch (ctx->mech_kind) {
        case CRM_MECHANICS_NEWTON_DAMPED:
            hipLaunchKernelGGL((k_gather_integrate_point<D, MIXED, CRM_MECHANICS_NEWTON_DAMPED>), dim3(ctx->n_work),
                               dim3(kBlock), 0, ctx->stream, A, src, dst, sa);
            return CRM_OK;
        case CRM_MECHANICS_LANGEVIN:
            hipLaunchKernelGGL((k_gather_integrate_point<D, MIXED, CRM_MECHANICS_LANGEVIN>), dim3(ctx->n_work),
                               dim3(kBlock), 0, ctx->stream, A, src, dst, sa);
            return CRM_OK;
        case CRM_MECHANICS_BROWNIAN:
            hipLaunchKernelGGL((k_gather_integrate_point<D, MIXED, CRM_MECHANICS_BROWNIAN>), dim3(ctx->n_work),
                               dim3(kBlock), 0, ctx->stream, A, src, dst, sa);
            return CRM_OK;
    }
    return CRM_INVALID_ARGUMENT;
}

// Initial placement gather; NewtonDamped / Langevin also bin the first step.
template <int D, bool MIXED>
crm_status launch_gather_setup_d(crm_ctx* ctx, StateBuf src, StateBuf dst, const StepArgs& sa) {
    Arrays A = arrays(ctx);
    const dim3 g(ctx->n_work), b(kBlock);
    switch (ctx->mech_kind) {
        case CRM_MECHANICS_NEWTON_DAMPED:
            hipLaunchKernelGGL((k_gather_point<D, MIXED, true, CRM_MECHANICS_NEWTON_DAMPED>), g, b, 0, ctx->stream, A,
                               src, dst, sa);
            return CRM_OK;
        case CRM_MECHANICS_LANGEVIN:
            hipLaunchKernelGGL((k_gather_point<D, MIXED, true, CRM_MECHANICS_LANGEVIN>), g, b, 0, ctx->stream, A, src,
                               dst, sa);
            return CRM_OK;
        case CRM_MECHANICS_BROWNIAN:
            hipLaunchKernelGGL((k_gather_point<D, MIXED, false, CRM_MECHANICS_BROWNIAN>), g, b, 0, ctx->stream, A,
                               src, dst, sa);
            return CRM_OK;
    }
    return CRM_INVALID_ARGUMENT;
}

bool keys_ahead(const crm_ctx* ctx);

// scan + scatter + fixup + gather: the counting sort that rebuilds the voxel
// lists from d_key / d_rank / d_count and moves the state to the other buffer.
crm_status launch_sort(crm_ctx* ctx, bool advance, const StepArgs* sa) {
    Arrays A = arrays(ctx);
    const uint32_t Vt = (uint32_t)ctx->Vt;
    if (ctx->n_tiles == 1 && ctx->N <= kSmallSortAgents) {  // one workgroup does all three (timed as scan)
        CRM_TRY(t_begin(ctx));
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(kBlock), 0, ctx->stream, A, (uint32_t)(advance ? 1 : 0));
        CRM_TRY(t_end(ctx, CRM_KERNEL_SCAN));
    } else {
        CRM_TRY(t_begin(ctx));
        hipLaunchKernelGGL(k_scan_local, dim3(ctx->n_tiles), dim3(kBlock), 0, ctx->stream, ctx->d_count,
                           ctx->d_count + stay_offset(ctx), ctx->d_start_local, ctx->d_bsum, Vt, ctx->d_ctrl);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_bsum, ctx->n_tiles,
                           ctx->d_start, Vt, ctx->d_ctrl, (uint32_t)(advance ? 1 : 0));
        CRM_TRY(t_end(ctx, CRM_KERNEL_SCAN));
        CRM_TRY(t_begin(ctx));
        const uint64_t n_scatter = std::max<uint64_t>(ctx->N, Vt);
        hipLaunchKernelGGL(k_scatter, dim3((uint32_t)((n_scatter + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, A);
        CRM_TRY(t_end(ctx, CRM_KERNEL_SCATTER));
    }
    // the point gather of a step orders each voxel's arrivals itself
    if (is_rod(ctx) || !advance) {
        CRM_TRY(t_begin(ctx));
        hipLaunchKernelGGL(k_fixup, dim3((Vt + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, A);
        CRM_TRY(t_end(ctx, CRM_KERNEL_FIXUP));
    }
    CRM_TRY(t_begin(ctx));
    StateBuf src = ctx->buf[ctx->cur], dst = ctx->buf[1 - ctx->cur];
    const uint32_t cap = ctx->cfg.solver_order;
    const bool mixed = ctx->cfg.precision == CRM_PRECISION_MIXED;
    crm_status gst = CRM_OK;
    if (is_rod(ctx)) {
        const RodArgs R = rod_args(ctx);
        const StepArgs sa0 = sa ? *sa : step_args(ctx, nullptr);
        const dim3 g(ctx->n_work), b(kBlock);
        const bool ahead = keys_ahead(ctx);
        if (ctx->D == 2) {
            if (ahead) {
                if (advance) hipLaunchKernelGGL((k_gather_rod<2, true, true>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
                else hipLaunchKernelGGL((k_gather_rod<2, false, true>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
            } else {
                if (advance) hipLaunchKernelGGL((k_gather_rod<2, true, false>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
                else hipLaunchKernelGGL((k_gather_rod<2, false, false>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
            }
        } else {
            if (ahead) {
                if (advance) hipLaunchKernelGGL((k_gather_rod<3, true, true>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
                else hipLaunchKernelGGL((k_gather_rod<3, false, true>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
            } else {
                if (advance) hipLaunchKernelGGL((k_gather_rod<3, true, false>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
                else hipLaunchKernelGGL((k_gather_rod<3, false, false>), g, b, 0, ctx->stream, A, src, dst, R, sa0);
            }
        }
    } else if (advance) {
        if (ctx->D == 2)
            gst = mixed ? launch_gather_integrate_dm<2, true>(ctx, src, dst, *sa)
                        : launch_gather_integrate_dm<2, false>(ctx, src, dst, *sa);
        else
            gst = mixed ? launch_gather_integrate_dm<3, true>(ctx, src, dst, *sa)
                        : launch_gather_integrate_dm<3, false>(ctx, src, dst, *sa);
    } else {
        const StepArgs sa0 = sa ? *sa : step_args(ctx, nullptr);
        gst = ctx->D == 2 ? (mixed ? launch_gather_setup_d<2, true>(ctx, src, dst, sa0)
                                   : launch_gather_setup_d<2, false>(ctx, src, dst, sa0))
                          : (mixed ? launch_gather_setup_d<3, true>(ctx, src, dst, sa0)
                                   : launch_gather_setup_d<3, false>(ctx, src, dst, sa0));
    }
    if (gst) return gst;
    CRM_TRY(t_end(ctx, CRM_KERNEL_GATHER));
    ctx->cur = 1 - ctx->cur;
    return CRM_OK;
}

// NewtonDamped and Langevin point agents, and rods outside the cycle, are binned
// one step ahead by the gather (their position update needs no force, and no
// noise but Philox's of the next step, so not with a caller's noise buffer); the
// others get a key kernel after the forces.
bool keys_ahead(const crm_ctx* ctx) {
    if (is_rod(ctx)) return !ctx->cycle_on && ctx->cfg.noise_mode != CRM_NOISE_BUFFER;
    return ctx->mech_kind != CRM_MECHANICS_BROWNIAN;
}

crm_status launch_step(crm_ctx* ctx, const double* noise_step) {
    const StepArgs sa = step_args(ctx, noise_step);
    const bool fused_rank = keys_ahead(ctx) && !is_rod(ctx);  // the force tiles rank their voxels
    CRM_TRY(launch_force(ctx, fused_rank));
    if (!keys_ahead(ctx)) {
        CRM_TRY(launch_key(ctx, sa));
    } else if (!fused_rank) {
        CRM_TRY(t_begin(ctx));
        hipLaunchKernelGGL(k_rank_ahead, dim3((uint32_t)((ctx->Vt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, arrays(ctx));
        CRM_TRY(t_end(ctx, CRM_KERNEL_INTEGRATE));
    }
    CRM_TRY(launch_sort(ctx, true, &sa));
    return CRM_OK;
}

crm_status read_control(crm_ctx* ctx, Control& c) {
    CRM_HIP(hipMemcpyAsync(&c, ctx->d_ctrl, sizeof(Control), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

// Maps the device error word to a status and message.
crm_status check_errors(crm_ctx* ctx, const char* phase) {
    Control c;
    CRM_TRY(read_control(ctx, c));
    if (!c.error_flags) return CRM_OK;
    crm_status kind = CRM_CALC_ERROR;
    const char* msg = "CalcError: identical position for two objects. Cannot Calculate force in Mie potential";
    if (c.error_flags & (1u << CRM_CALC_ERROR)) {
    } else if (c.error_flags & (1u << CRM_BOUNDARY_ERROR)) {
        kind = CRM_BOUNDARY_ERROR;
        msg = "BoundaryError: Particle is out of domain";
    } else if (c.error_flags & (1u << CRM_INDEX_ERROR)) {
        kind = CRM_INDEX_ERROR;
        msg = "DecomposeError: could not find voxel index";
    } else if (c.error_flags & (1u << CRM_DIVISION_ERROR)) {
        kind = CRM_DIVISION_ERROR;
        msg = "DivisionError: Could not calculate new segment lengths";
    } else {
        kind = CRM_DEVICE_ERROR;
        msg = "unknown device error";
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s (%s, step %llu, agent %llu)", msg, phase,
                  (unsigned long long)(c.error_key >> 32), (unsigned long long)(c.error_key & 0xffffffffu));
    ctx->err = buf;
    ctx->halted = true;
    return kind;
}

crm_status reset_control(crm_ctx* ctx, uint64_t step) {
    Control c;
    std::memset(&c, 0, sizeof(c));
    c.step = step;
    c.error_key = ~0ull;
    CRM_HIP(hipMemcpyAsync(ctx->d_ctrl, &c, sizeof(Control), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

// Force tiles of one instance (see k_force_tiled): whole rows when the expected
// neighbourhood of a few rows fits the LDS budget, otherwise single rows cut
// along their length.  Sized from the mean density with a 20 % margin; a tile
// that overflows at run time falls back to global loads.
// Force tiles of one instance: plane a, R rows x W columns of voxels each, the
// grid split evenly.  The shape keeps the expected own agents within one pass of
// the workgroup (lanes busy), the expected staged region (x1.2 for density
// fluctuations) within the LDS records, the staged voxels and the own voxels
// within their LDS tables and the region extent within 31 voxels (>= 26
// fixed-point bits per voxel); among those it takes the most own agents, then
// the smallest region.
// Tiling policies the autotuner chooses from: the fraction of the workgroup's
// lanes the expected own agents may fill, and the density margin on the staged
// region.  Policy 0 is the default.
// own_frac: the tile's expected own agents per lane of the workgroup (above 1, a
// lane walks several own agents one after another, which fills the lanes of
// tiles cut from small grids)
struct TilePolicy {
    double own_frac, margin;
    uint32_t lanes_log2;  // lanes per own agent: 1, 2, 4 or 8 (agents per wavefront 64, 32, 16, 8)
};
constexpr TilePolicy kTilePolicies[] = {{0.94, 1.2, 0}, {0.94, 1.1, 0}, {0.94, 1.4, 0}, {0.80, 1.2, 0},
                                        {0.70, 1.2, 0}, {0.55, 1.2, 0}, {0.40, 1.2, 0}, {0.70, 1.4, 0},
                                        {0.94, 1.2, 1}, {0.70, 1.2, 1}, {0.94, 1.2, 2}, {0.70, 1.2, 2},
                                        {1.50, 1.2, 0}, {1.90, 1.2, 0}, {0.94, 1.2, 3}, {0.70, 1.2, 3}};
constexpr uint32_t kNumTilePolicies = sizeof(kTilePolicies) / sizeof(kTilePolicies[0]);

void build_tiles(const InstParams& P, uint32_t dim, uint32_t inst, uint64_t n_agents, uint32_t cap,
                 const TilePolicy& pol, std::vector<ForceTile>& out, uint32_t& rows, uint32_t& cols_out) {
    const uint32_t n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    const double rho = (double)n_agents / ((double)n0 * n1 * n2);
    const uint32_t np = dim == 3 ? 3u : 1u;  // staged planes
    const double planes = (double)std::min(np, n0 > 1 ? 3u : 1u);
    const double margin = pol.margin, own_max = pol.own_frac * kBlock / (double)(1u << pol.lanes_log2);
    uint32_t bR = 1, bW = 1;
    double best_own = -1.0, best_staged = 0.0;
    for (uint32_t r0 = 1; r0 <= std::min<uint32_t>(n1, kTileRowsMax); r0++) {
        const uint32_t nb = (n1 + r0 - 1) / r0, R = (n1 + nb - 1) / nb;
        for (uint32_t w0 = 1; w0 <= std::min<uint32_t>(n2, 29u); w0++) {
            const uint32_t nc = (n2 + w0 - 1) / w0, W = (n2 + nc - 1) / nc;
            const uint32_t cols = std::min(W + 2, n2);
            const double own = (double)R * W * rho;
            const double staged = planes * (R + 2) * (W + 2) * rho * margin;
            if (own > own_max || staged > cap) continue;
            if (np * (R + 2) * cols + 1 > (uint32_t)kLdsVoxelStarts || R * W > (uint32_t)kOwnVoxMax ||
                std::max(R + 2, cols) > 31u)
                continue;
            if (own > best_own * (1.0 + 1e-9) || (own >= best_own * (1.0 - 1e-9) && staged < best_staged)) {
                best_own = own;
                best_staged = staged;
                bR = R;
                bW = W;
            }
        }
    }
    for (uint32_t a = 0; a < n0; a++)
        for (uint32_t b = 0; b < n1; b += bR)
            for (uint32_t c = 0; c < n2; c += bW)
                out.push_back(ForceTile{inst, a, b, std::min(b + bR, n1), c, std::min(c + bW, n2)});
    rows = bR;
    cols_out = bW;
}

template <typename T>
crm_status dalloc(crm_ctx* ctx, T*& p, size_t n) {
    CRM_HIP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return CRM_OK;
}

// Force tiles of every instance under tiling policy `pol`.
void make_tiles(crm_ctx* ctx, uint32_t pol, std::vector<ForceTile>& tiles) {
    const uint32_t cap = ctx->cfg.precision == CRM_PRECISION_MIXED ? kLdsRecordsMixed : kLdsRecordsFP64;
    tiles.clear();
    ctx->tile_rows.assign(ctx->n_inst, 0);
    ctx->tile_cols.assign(ctx->n_inst, 0);
    ctx->tile_count.assign(ctx->n_inst, 0);
    for (uint32_t i = 0; i < ctx->n_inst; i++) {
        const size_t before = tiles.size();
        build_tiles(ctx->hparams[i], ctx->D, i, ctx->counts[i], cap, kTilePolicies[pol], tiles, ctx->tile_rows[i],
                    ctx->tile_cols[i]);
        ctx->tile_count[i] = (uint32_t)(tiles.size() - before);
    }
    ctx->lanes_log2 = kTilePolicies[pol].lanes_log2;
}

crm_status upload_tiles(crm_ctx* ctx, const std::vector<ForceTile>& tiles) {
    if (tiles.size() > ctx->tiles_cap) {
        if (ctx->d_tiles) CRM_HIP(hipFree(ctx->d_tiles));
        ctx->d_tiles = nullptr;
        CRM_TRY(dalloc(ctx, ctx->d_tiles, tiles.size()));
        ctx->tiles_cap = tiles.size();
    }
    CRM_HIP(hipMemcpy(ctx->d_tiles, tiles.data(), sizeof(ForceTile) * tiles.size(), hipMemcpyHostToDevice));
    ctx->n_force_tiles = (uint32_t)tiles.size();
    return CRM_OK;
}

// Downloads a [C][N] SoA array and reorders it to agent index (rows of C).
crm_status download_by_id(crm_ctx* ctx, const double* d_src, uint32_t C, double* out,
                          const std::vector<uint32_t>& id) {
    std::vector<double> tmp((size_t)C * ctx->N);
    CRM_HIP(hipMemcpyAsync(tmp.data(), d_src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t s = 0; s < ctx->N; s++)
        for (uint32_t k = 0; k < C; k++) out[(uint64_t)id[s] * C + k] = tmp[(uint64_t)k * ctx->N + s];
    return CRM_OK;
}

crm_status download_ids(crm_ctx* ctx, std::vector<uint32_t>& id) {
    id.resize(ctx->N);
    CRM_HIP(hipMemcpyAsync(id.data(), ctx->buf[ctx->cur].id, ctx->N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

crm_status build_graph(crm_ctx* ctx, int parity, bool long_graph) {
    // An even number of steps returns the buffers to the same parity, so one
    // captured sequence of two (or kLongGraphSteps) steps replays from either parity.
    int saved = ctx->cur;
    ctx->cur = parity;
    CRM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    crm_status st = CRM_OK;
    for (uint32_t k = 0; k < (long_graph ? kLongGraphSteps : 2u) && st == CRM_OK; k++) st = launch_step(ctx, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->cur = saved;
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        ctx->err = std::string("HIP error: ") + hipGetErrorString(e) + " at hipStreamEndCapture";
        return CRM_DEVICE_ERROR;
    }
    e = hipGraphInstantiate(&ctx->graph[(long_graph ? 2 : 0) + parity], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        ctx->err = std::string("HIP error: ") + hipGetErrorString(e) + " at hipGraphInstantiate";
        return CRM_DEVICE_ERROR;
    }
    return CRM_OK;
}

// Temporary storage of the daughters' exclusive scan (hipcub, Vt + 1 items).
crm_status alloc_scan_tmp(crm_ctx* ctx) {
    size_t bytes = 0;
    CRM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, ctx->d_ndiv, ctx->d_dbefore, (int)(ctx->Vt + 1),
                                             ctx->stream));
    uint8_t* p = nullptr;
    CRM_TRY(dalloc(ctx, p, std::max<size_t>(bytes, 1)));
    ctx->d_scan_tmp = p;
    ctx->scan_tmp_bytes = std::max<size_t>(bytes, 1);
    return CRM_OK;
}

// Work blocks (rods_per_block rods, never across instances) for the current
// per-instance counts; offsets and counts go to the instance parameters.  The
// uploads are queued on the stream (the host copies live in the context).
crm_status rebuild_work(crm_ctx* ctx) {
    std::vector<BlockWork>& work = ctx->work_host;
    work.clear();
    uint64_t first = 0;
    const uint64_t per = is_rod(ctx) ? ctx->rods_per_block : (uint64_t)kBlock;
    for (uint32_t i = 0; i < ctx->n_inst; i++) {
        ctx->hparams[i].agent_offset = first;
        ctx->hparams[i].n_agents = ctx->counts[i];
        for (uint64_t b = 0; b < ctx->counts[i]; b += per) {
            BlockWork w;
            w.inst = i;
            w.first = first + b;
            w.count = (uint32_t)std::min<uint64_t>(per, ctx->counts[i] - b);
            work.push_back(w);
        }
        first += ctx->counts[i];
    }
    ctx->n_work = (uint32_t)work.size();
    CRM_HIP(hipMemcpyAsync(ctx->d_work, work.data(), sizeof(BlockWork) * work.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    CRM_HIP(hipMemcpyAsync(ctx->d_ip, ctx->hparams.data(), sizeof(InstParams) * ctx->n_inst, hipMemcpyHostToDevice,
                           ctx->stream));
    return CRM_OK;
}

// One step with the rods example's cycle (kernels_cycle.hpp).  The host reads the
// control block once per step (errors so far, divisions per instance) to size the
// grown order; everything else stays queued on the stream.
crm_status launch_cycle_step(crm_ctx* ctx, const double* noise_step) {
    const StepArgs sa = step_args(ctx, noise_step);
    const uint32_t D = ctx->D, Vt = (uint32_t)ctx->Vt;
    hipLaunchKernelGGL(k_cycle_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->d_ctrl, ctx->d_inst_div, ctx->n_inst);
    CRM_TRY(launch_force(ctx));
    CycleArgs Y;
    Y.base = ctx->cyc.growth_base_rate;
    Y.threshold = ctx->cyc.division_threshold;
    Y.saturation = ctx->cyc.neighbor_saturation;
    Y.divflag = ctx->d_divflag;
    Y.drank = ctx->d_drank;
    Y.dctr = ctx->d_dctr;
    Y.ndiv = ctx->d_ndiv;
    Y.dbefore = ctx->d_dbefore;
    Y.vcount = ctx->d_vcount;
    Y.inst_div = ctx->d_inst_div;
    Y.n_inst = ctx->n_inst;
    const RodArgs R = rod_args(ctx);
    StateBuf B = ctx->buf[ctx->cur], B2 = ctx->buf[1 - ctx->cur];
    {
        Arrays A = arrays(ctx);
        const dim3 g(ctx->n_work), b(kBlock);
        if (D == 2) hipLaunchKernelGGL(k_rod_mech<2>, g, b, 0, ctx->stream, A, B, B2, R, sa, Y);
        else hipLaunchKernelGGL(k_rod_mech<3>, g, b, 0, ctx->stream, A, B, B2, R, sa, Y);
        CRM_HIP(hipGetLastError());
    }
    Control c;
    std::vector<uint32_t> inst_div(ctx->n_inst);
    CRM_HIP(hipMemcpyAsync(&c, ctx->d_ctrl, sizeof(Control), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(inst_div.data(), ctx->d_inst_div, ctx->n_inst * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    // an error of this step lets the step end (as the plain step does); one of
    // the step before halted this step's kernels
    if (c.error_flags && (c.error_key >> 32) < ctx->step) return check_errors(ctx, "step");
    const uint64_t n_old = ctx->N, n_new = n_old + c.n_div;
    if (c.n_div == 0) {
        ctx->cur = 1 - ctx->cur;  // the updated state is in the other buffer
    } else {
        if (n_new > ctx->cap_agents) {
            ctx->err = "the cycle needs " + std::to_string(n_new) + " rods but the context holds max_agents = " +
                       std::to_string(ctx->cap_agents);
            ctx->halted = true;
            return CRM_INVALID_ARGUMENT;
        }
        Arrays A = arrays(ctx);
        hipLaunchKernelGGL(k_cycle_voxels, dim3((Vt + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, A,
                           B2, Y);
        CRM_HIP(hipGetLastError());
        size_t bytes = ctx->scan_tmp_bytes;
        CRM_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->d_scan_tmp, bytes, ctx->d_ndiv, ctx->d_dbefore, (int)(Vt + 1),
                                                 ctx->stream));
        const dim3 g((uint32_t)((n_old + kBlock - 1) / kBlock)), b(kBlock);
        if (D == 2)
            hipLaunchKernelGGL(k_expand_rod<2>, g, b, 0, ctx->stream, A, B2, B, R, sa, Y, n_old, n_new,
                               ctx->cyc.division_radius);
        else
            hipLaunchKernelGGL(k_expand_rod<3>, g, b, 0, ctx->stream, A, B2, B, R, sa, Y, n_old, n_new,
                               ctx->cyc.division_radius);
        CRM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_grow_start, dim3((Vt + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream,
                           ctx->d_start, ctx->d_dbefore, Vt);
        CRM_HIP(hipGetLastError());
        for (uint32_t i = 0; i < ctx->n_inst; i++) ctx->counts[i] += inst_div[i];
        ctx->N = n_new;
        CRM_TRY(rebuild_work(ctx));
    }
    {
        Arrays A = arrays(ctx);
        const dim3 g(ctx->n_work), b(kBlock);
        if (D == 2) hipLaunchKernelGGL(k_reflect_key_rod<2>, g, b, 0, ctx->stream, A, ctx->buf[ctx->cur], R);
        else hipLaunchKernelGGL(k_reflect_key_rod<3>, g, b, 0, ctx->stream, A, ctx->buf[ctx->cur], R);
        CRM_HIP(hipGetLastError());
    }
    CRM_TRY(launch_rank(ctx));
    CRM_TRY(launch_sort(ctx, false, &sa));
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl);
    CRM_HIP(hipGetLastError());
    return CRM_OK;
}

// History entries stored after ctx->step steps and their ring slots, oldest
// first (the ring slot of push t is t % capacity).
uint32_t history_slots(const crm_ctx* ctx, uint32_t slot[2]) {
    const uint32_t cap = ctx->cfg.solver_order;
    const uint64_t t = ctx->step;
    const uint32_t size = cap == 0 ? 0u : (uint32_t)std::min<uint64_t>(t, cap);
    const uint32_t oldest = (size < cap || cap == 0) ? 0u : (uint32_t)(t % cap);
    for (uint32_t i = 0; i < 2; i++) slot[i] = cap ? (oldest + i) % cap : 0u;
    return size;
}

// Gathers the snapshot records (kernels_snapshot.hpp) into agent order in
// d_snap / d_snap_sp; returns their length in doubles.
crm_status gather_snapshot_device(crm_ctx* ctx, uint32_t& n_hist, size_t& len_out) {
    SnapshotArgs a{};
    a.N = ctx->N;
    a.K = ctx->K;
    a.n_hist = history_slots(ctx, a.slot);
    n_hist = a.n_hist;
    const size_t len = (size_t)ctx->N * ctx->K * (2u + 2u * a.n_hist);
    const bool held = is_rod(ctx) || ctx->pt_births;  // agents hold rings of different fill
    if (len > ctx->snap_cap || (held && !ctx->d_snap_st)) {
        void* old[] = {ctx->d_snap, ctx->d_snap_sp, ctx->d_snap_rod, ctx->d_snap_id, ctx->d_snap_st};
        for (void* p : old)
            if (p) CRM_HIP(hipFree(p));
        ctx->d_snap = nullptr;
        ctx->d_snap_sp = nullptr;
        ctx->d_snap_rod = nullptr;
        ctx->d_snap_id = nullptr;
        ctx->d_snap_st = nullptr;
        ctx->snap_cap = 0;
        CRM_TRY(dalloc(ctx, ctx->d_snap, len));
        CRM_TRY(dalloc(ctx, ctx->d_snap_sp, ctx->N));
        if (is_rod(ctx)) {
            CRM_TRY(dalloc(ctx, ctx->d_snap_rod, 2 * ctx->N));
            CRM_TRY(dalloc(ctx, ctx->d_snap_id, 6 * ctx->N));
        }
        if (held) CRM_TRY(dalloc(ctx, ctx->d_snap_st, ctx->N));
        ctx->snap_cap = len;
    }
    a.out = ctx->d_snap;
    a.species = ctx->d_snap_sp;
    a.rod = ctx->d_snap_rod;
    a.ident = ctx->d_snap_id;
    a.stored = held ? ctx->d_snap_st : nullptr;
    a.step = ctx->step;
    hipLaunchKernelGGL(k_snapshot, dim3((uint32_t)((ctx->N + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->buf[ctx->cur], a);
    CRM_HIP(hipGetLastError());
    len_out = len;
    return CRM_OK;
}

// The same, copied to host vectors.
crm_status gather_snapshot(crm_ctx* ctx, std::vector<double>& rec, std::vector<int32_t>& species, uint32_t& n_hist) {
    size_t len = 0;
    CRM_TRY(gather_snapshot_device(ctx, n_hist, len));
    rec.resize(len);
    species.resize(ctx->N);
    CRM_HIP(hipMemcpyAsync(rec.data(), ctx->d_snap, len * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(species.data(), ctx->d_snap_sp, ctx->N * sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

// A double as serde_json writes an f64: shortest round-trip digits, a fraction
// or an exponent always present, the exponent without '+' or leading zeros;
// non-finite values become null.
void put_num(std::string& o, double x) {
    if (!std::isfinite(x)) {
        o += "null";
        return;
    }
    char b[48];
    const auto r = std::to_chars(b, b + sizeof b, x);
    const char* e = std::find(b, r.ptr, 'e');
    if (e == r.ptr) {
        o.append(b, (size_t)(r.ptr - b));
        if (std::find(b, r.ptr, '.') == r.ptr) o += ".0";
        return;
    }
    o.append(b, (size_t)(e - b));
    o += 'e';
    const char* q = e + 1;
    if (*q == '+') q++;
    else if (*q == '-') o += *q++;
    while (q + 1 < r.ptr && *q == '0') q++;
    o.append(q, (size_t)(r.ptr - q));
}

// A position-like value: nalgebra SVector (points) serialises as [x, y, z];
// the rods' MatrixXx{D} as [column-major data, rows, columns].
void put_vec(std::string& o, const double* v, uint32_t nv, uint32_t D, bool matrix) {
    if (!matrix) {
        o += '[';
        for (uint32_t k = 0; k < D; k++) {
            if (k) o += ", ";
            put_num(o, v[k]);
        }
        o += ']';
        return;
    }
    o += "[[";
    for (uint32_t k = 0; k < D; k++)
        for (uint32_t i = 0; i < nv; i++) {
            if (k || i) o += ", ";
            put_num(o, v[i * D + k]);
        }
    o += "], " + std::to_string(nv) + ", " + std::to_string(D) + "]";
}

// A cell template split at its placeholders {pos}, {vel}, {spring_length},
// {growth_rate} (the last two: the rod's values).
struct Template {
    std::vector<std::string> lit;  // lit[0] hole[0] lit[1] hole[1] ... lit[n]
    std::vector<int> hole;         // 0 pos, 1 vel, 2 spring length, 3 growth rate
};

Template split_template(const char* t) {
    static const char* names[4] = {"{pos}", "{vel}", "{spring_length}", "{growth_rate}"};
    Template T;
    std::string cur;
    for (const char* p = t; *p;) {
        int h = -1;
        for (int q = 0; q < 4 && h < 0; q++)
            if (std::strncmp(p, names[q], std::strlen(names[q])) == 0) h = q;
        if (h >= 0) {
            T.lit.push_back(cur);
            cur.clear();
            T.hole.push_back(h);
            p += std::strlen(names[h]);
        } else {
            cur += *p++;
        }
    }
    T.lit.push_back(cur);
    return T;
}

// CellIdentifier as serde writes the externally tagged enum: {"Initial": a},
// {"Division": [a, b]}; kind 3 (no parent) is null.
std::string ident_json(const uint64_t* x) {
    if (x[0] == 0) return "{\"Initial\": " + std::to_string(x[1]) + "}";
    if (x[0] == 1) return "{\"Division\": [" + std::to_string(x[1]) + ", " + std::to_string(x[2]) + "]}";
    if (x[0] == 2) return "{\"Inserted\": [" + std::to_string(x[1]) + ", " + std::to_string(x[2]) + "]}";
    return "null";
}

// One data entry of the batch file: CombinedSaveFormat { identifier, element:
// (CellBox { identifier, parent, cell }, AuxStorage { mechanics }) }
// (storage/concepts.rs:176-190, cell.rs:128-137, aux_storage.rs:41-46).
// ids [6] (identifier, parent); rods: rv [2] (spring length, growth rate); the
// history entries the agent holds are the newest `held` of the n_hist slots.
void put_cell(std::string& o, const double* r, const Template& T, uint32_t nv, uint32_t D, uint32_t n_hist,
              const uint64_t* ids, const double* rv, uint32_t held) {
    const uint32_t K = nv * D;
    const bool matrix = nv > 1;
    const std::string ident = ident_json(ids), parent = ident_json(ids + 3);
    o += "    {\n      \"identifier\": " + ident + ",\n      \"element\": [\n        {\n          \"identifier\": " +
         ident + ",\n          \"parent\": " + parent + ",\n          \"cell\": ";
    for (size_t i = 0; i < T.hole.size(); i++) {
        o += T.lit[i];
        const int h = T.hole[i];
        if (h <= 1) put_vec(o, r + (h == 0 ? 0u : K), nv, D, matrix);
        else put_num(o, rv ? rv[h - 2] : 0.0);
    }
    o += T.lit.back();
    o += "\n        },\n        {\n          \"mechanics\": {\n            \"positions\": [";
    const uint32_t first = n_hist - std::min(held, n_hist);
    for (uint32_t i = first; i < n_hist; i++) {
        if (i > first) o += ", ";
        put_vec(o, r + 2u * K + i * K, nv, D, matrix);
    }
    o += "],\n            \"velocities\": [";
    for (uint32_t i = first; i < n_hist; i++) {
        if (i > first) o += ", ";
        put_vec(o, r + 2u * K + (n_hist + i) * K, nv, D, matrix);
    }
    std::vector<double> zero(K, 0.0);
    o += "],\n            \"current_force\": ";
    put_vec(o, zero.data(), nv, D, matrix);
    o += ",\n            \"zero_force\": ";
    put_vec(o, zero.data(), nv, D, matrix);
    o += "\n          }\n        }\n      ]\n    }";
}

// Checks of a removal call; n_rm = agents flagged.
crm_status remove_prelude(crm_ctx* ctx, const uint8_t* remove, uint64_t& n_rm) {
    if (!ctx || !remove) return CRM_INVALID_ARGUMENT;
    if (!ctx->have_agents) {
        ctx->err = "crm_set_agents must be called first";
        return CRM_INVALID_ARGUMENT;
    }
    if (ctx->halted) return CRM_INVALID_ARGUMENT;
    n_rm = 0;
    for (uint64_t a = 0; a < ctx->N; a++) n_rm += remove[a] ? 1u : 0u;
    if (n_rm == ctx->N) {
        ctx->err = "the removal would leave no agents";
        return CRM_INVALID_ARGUMENT;
    }
    return CRM_OK;
}

// Removal scratch for n_old agents, carved at 256-byte aligned offsets: flags
// [n_old] bytes, four [n_old + 1] uint32 arrays, [n_inst] uint32, the scan's storage.
struct RemoveLayout {
    size_t keep_slot, keep_id, slot_before, id_before, inst, scan, scan_bytes, total;
};

crm_status remove_layout(crm_ctx* ctx, uint64_t n_old, RemoveLayout& L) {
    const size_t n1 = n_old + 1;
    size_t bytes = 0;
    CRM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n1,
                                             ctx->stream));
    L.keep_slot = align256(n_old);
    L.keep_id = L.keep_slot + align256(4 * n1);
    L.slot_before = L.keep_id + align256(4 * n1);
    L.id_before = L.slot_before + align256(4 * n1);
    L.inst = L.id_before + align256(4 * n1);
    L.scan = L.inst + align256(4 * (size_t)ctx->n_inst);
    L.scan_bytes = std::max<size_t>(bytes, 1);
    L.total = L.scan + align256(L.scan_bytes);
    return CRM_OK;
}

// CycleEvent::Remove on the device with the scratch at `base` (layout L): survivor
// flags, their scans, compaction into the other state buffer, then a set-up sort.
// Allocates no device memory.
crm_status remove_agents_impl(crm_ctx* ctx, const uint8_t* remove, uint64_t n_rm, uint8_t* base,
                              const RemoveLayout& L) {
    const uint64_t n_old = ctx->N, n_new = n_old - n_rm;
    CRM_HIP(hipSetDevice(ctx->device));
    RemoveArgs X{};
    X.flag = base;
    X.keep_slot = reinterpret_cast<uint32_t*>(base + L.keep_slot);
    X.keep_id = reinterpret_cast<uint32_t*>(base + L.keep_id);
    X.slot_before = reinterpret_cast<uint32_t*>(base + L.slot_before);
    X.id_before = reinterpret_cast<uint32_t*>(base + L.id_before);
    X.inst_rm = reinterpret_cast<uint32_t*>(base + L.inst);
    void* scan_tmp = base + L.scan;
    size_t bytes = L.scan_bytes;
    X.n_old = n_old;
    X.n_inst = ctx->n_inst;
    CRM_HIP(hipMemcpyAsync(base, remove, n_old, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3((ctx->n_inst + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream,
                       X.inst_rm, ctx->n_inst, 0u);
    const StateBuf src = ctx->buf[ctx->cur], dst = ctx->buf[1 - ctx->cur];
    hipLaunchKernelGGL(k_remove_mark, dim3((uint32_t)((n_old + 1 + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, arrays(ctx), src, X);
    CRM_HIP(hipGetLastError());
    CRM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, X.keep_slot, X.slot_before, (int)(n_old + 1),
                                             ctx->stream));
    CRM_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, X.keep_id, X.id_before, (int)(n_old + 1),
                                             ctx->stream));
    hipLaunchKernelGGL(k_remove_compact, dim3((uint32_t)((n_old + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, src, dst, X, n_new, ctx->K, ctx->cfg.solver_order, (uint32_t)(is_rod(ctx) ? 1 : 0));
    CRM_HIP(hipGetLastError());
    uint32_t* inst_rm = ctx->inst_host.data();
    CRM_HIP(hipMemcpyAsync(inst_rm, X.inst_rm, sizeof(uint32_t) * ctx->n_inst, hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t total = 0;
    for (uint32_t i = 0; i < ctx->n_inst; i++) {
        if (inst_rm[i] > ctx->counts[i]) total = UINT64_MAX;
        if (total != UINT64_MAX) total += inst_rm[i];
    }
    if (total != n_rm) {
        ctx->err = "agent removal: the removed agents do not match the instances' voxels";
        ctx->halted = true;
        return CRM_DEVICE_ERROR;
    }
    for (uint32_t i = 0; i < ctx->n_inst; i++) ctx->counts[i] -= inst_rm[i];
    if (!is_rod(ctx)) {  // identifiers follow the agent indices, in order (shrinks in place)
        uint64_t m = 0;
        for (uint64_t a = 0; a < n_old; a++) {
            if (remove[a]) continue;
            for (int q = 0; q < 3; q++) {
                ctx->pt_ident[3 * m + q] = ctx->pt_ident[3 * a + q];
                ctx->pt_parent[3 * m + q] = ctx->pt_parent[3 * a + q];
            }
            m++;
        }
        ctx->pt_ident.resize(3 * n_new);
        ctx->pt_parent.resize(3 * n_new);
    }
    ctx->N = n_new;
    ctx->cur = 1 - ctx->cur;
    CRM_TRY(rebuild_work(ctx));
    ctx->sense_parity = -1;
    ctx->graphs_stale = true;  // captured for the old agent count
    // the voxel lists of the survivors: a set-up sort with every agent in its voxel
    CRM_HIP(hipMemsetAsync(ctx->d_count, 0, sizeof(uint32_t) * (stay_offset(ctx) + ctx->Vt), ctx->stream));
    hipLaunchKernelGGL(k_remove_key, dim3((uint32_t)((n_new + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       arrays(ctx), ctx->buf[ctx->cur]);
    CRM_HIP(hipGetLastError());
    CRM_TRY(launch_sort(ctx, false, nullptr));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return check_errors(ctx, "agent removal");
}

// ---- insertion of divided point agents (crm_insert_agents) ------------------

// apply_boundary of one point (cartesian_cuboid_n.rs:594-631), the reflection
// point_step applies; true if a coordinate is still outside afterwards.
bool reflect_point(const InstParams& P, uint32_t D, double* x, double* v) {
    bool out = false;
    for (uint32_t k = 0; k < D; k++) {
        const uint32_t ax = k + 3 - D;
        if (x[k] < P.min[ax]) {
            x[k] = 2.0 * P.min[ax] - x[k];
            v[k] = std::fabs(v[k]);
        }
        if (x[k] > P.max[ax]) {
            x[k] = 2.0 * P.max[ax] - x[k];
            v[k] = -std::fabs(v[k]);
        }
        out = out || x[k] < P.min[ax] || x[k] > P.max[ax];
    }
    return out;
}

// The device's axis_index (cartesian_cuboid_n.rs:469-491) on the host: the same
// IEEE operations, so the same voxel.
bool axis_index_host(double p, double mn, double dx, uint32_t n, uint32_t& out) {
    const double x = (p - mn) / dx;
    if (!(x > -1.0) || !(x < 4294967296.0)) return false;
    const double t = std::trunc(x);
    const uint32_t i = t <= 0.0 ? 0u : (uint32_t)t;
    if (i >= n) return false;
    out = i;
    return true;
}

uint32_t instance_of_voxel(const crm_ctx* ctx, uint64_t g) {
    uint32_t lo = 0, hi = ctx->n_inst;  // largest i with voxel_offset <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (ctx->hparams[mid].voxel_offset <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Every agent-sized device array for C agents (points), the graphs captured on
// the old ones released first.  The state buffers get birth steps.
crm_status realloc_point_arrays(crm_ctx* ctx, uint64_t C) {
    for (auto& g : ctx->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    void* old[] = {ctx->d_work, ctx->d_key, ctx->d_rank, ctx->d_perm, ctx->d_force, ctx->d_recq, ctx->d_rec64,
                   ctx->d_sense, ctx->d_snap, ctx->d_snap_sp, ctx->d_snap_st};
    for (void* q : old)
        if (q) CRM_HIP(hipFree(q));
    ctx->d_work = nullptr;
    ctx->d_key = ctx->d_rank = ctx->d_perm = nullptr;
    ctx->d_force = nullptr;
    ctx->d_recq = nullptr;
    ctx->d_rec64 = nullptr;
    ctx->d_sense = nullptr;
    ctx->d_snap = nullptr;
    ctx->d_snap_sp = nullptr;
    ctx->d_snap_st = nullptr;
    ctx->snap_cap = 0;
    for (auto& b : ctx->buf) {
        void* q[] = {b.pos, b.vel, b.hdx, b.hdv, b.species, b.id, b.vox, b.birth};
        for (void* x : q)
            if (x) CRM_HIP(hipFree(x));
        b = StateBuf{};
    }
    const uint32_t K = ctx->K;
    CRM_TRY(dalloc(ctx, ctx->d_work, C / kBlock + ctx->n_inst + 1));
    CRM_TRY(dalloc(ctx, ctx->d_key, 2 * C));
    CRM_TRY(dalloc(ctx, ctx->d_rank, C));
    CRM_TRY(dalloc(ctx, ctx->d_perm, C));
    CRM_TRY(dalloc(ctx, ctx->d_force, (size_t)K * C));
    if (ctx->cfg.precision == CRM_PRECISION_MIXED) CRM_TRY(dalloc(ctx, ctx->d_recq, C));
    else CRM_TRY(dalloc(ctx, ctx->d_rec64, C));
    if (ctx->sense_on) CRM_TRY(dalloc(ctx, ctx->d_sense, C));
    for (auto& b : ctx->buf) {
        CRM_TRY(dalloc(ctx, b.pos, (size_t)K * C));
        CRM_TRY(dalloc(ctx, b.vel, (size_t)K * C));
        CRM_TRY(dalloc(ctx, b.hdx, (size_t)2 * K * C));
        CRM_TRY(dalloc(ctx, b.hdv, (size_t)2 * K * C));
        CRM_TRY(dalloc(ctx, b.species, C));
        CRM_TRY(dalloc(ctx, b.id, C));
        CRM_TRY(dalloc(ctx, b.vox, C));
        CRM_TRY(dalloc(ctx, b.birth, C));
    }
    ctx->cap_agents = C;
    return CRM_OK;
}

// CycleEvent::Division for point agents (update_cycle.rs:28-66), applied between
// steps.  Voxel by voxel in list order, each mother takes Division(voxel,
// ++id_counter) with its old identifier as parent and the caller's state after
// C::divide (history kept); then its voxel's daughters, in the mothers' order,
// join the end of the voxel's list with Division(voxel, ++id_counter), the
// mother's old identifier as parent and an empty AuxStorage.  Mothers and
// daughters then get apply_boundary and the re-binning of a step (stayers keep
// their order, arrivals follow by old voxel and old position).
crm_status insert_agents_impl(crm_ctx* ctx, uint64_t n, const uint64_t* mother, const double* mpos,
                              const double* mvel, const double* dpos, const double* dvel, const int32_t* dsp) {
    const uint64_t N = ctx->N, Vt = ctx->Vt, N2 = N + n;
    const uint32_t K = ctx->K, D = ctx->D;
    if (N2 >= 0xffffffffull) {
        ctx->err = "too many agents (more than 2^32 - 1)";
        return CRM_INVALID_ARGUMENT;
    }
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    const StateBuf S = ctx->buf[ctx->cur];
    std::vector<double> pos((size_t)K * N), vel((size_t)K * N), hdx((size_t)2 * K * N), hdv((size_t)2 * K * N);
    std::vector<int32_t> sp(N);
    std::vector<uint32_t> id(N), vox(N), birth(N, 0), start(Vt + 1);
    CRM_HIP(hipMemcpyAsync(pos.data(), S.pos, pos.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(vel.data(), S.vel, vel.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(hdx.data(), S.hdx, hdx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(hdv.data(), S.hdv, hdv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(sp.data(), S.species, N * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(id.data(), S.id, N * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(vox.data(), S.vox, N * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (S.birth)
        CRM_HIP(hipMemcpyAsync(birth.data(), S.birth, N * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipMemcpyAsync(start.data(), ctx->d_start, (Vt + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));

    std::vector<uint64_t> slot_of(N);
    for (uint64_t s = 0; s < N; s++) slot_of[id[s]] = s;
    std::vector<int64_t> mother_at(N, -1);  // slot -> call index
    for (uint64_t i = 0; i < n; i++) {
        if (mother[i] >= N) {
            ctx->err = "mother " + std::to_string(mother[i]) + " is not an agent index";
            return CRM_INVALID_ARGUMENT;
        }
        const uint64_t s = slot_of[mother[i]];
        if (mother_at[s] >= 0) {
            ctx->err = "agent " + std::to_string(mother[i]) + " divides twice";
            return CRM_INVALID_ARGUMENT;
        }
        mother_at[s] = (int64_t)i;
    }
    // the divided states, reflected into the domain (nothing changes on an error)
    std::vector<double> mp((size_t)K * n), mv((size_t)K * n), dp((size_t)K * n), dv((size_t)K * n);
    std::vector<int32_t> dspecies(n);
    const uint32_t nsp = ctx->inst[0].potential.n_species;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t s = slot_of[mother[i]];
        const InstParams& P = ctx->hparams[instance_of_voxel(ctx, vox[s])];
        for (uint32_t k = 0; k < K; k++) {
            mp[i * K + k] = mpos ? mpos[i * K + k] : pos[(uint64_t)k * N + s];
            mv[i * K + k] = mvel ? mvel[i * K + k] : vel[(uint64_t)k * N + s];
            dp[i * K + k] = dpos[i * K + k];
        }
        for (uint32_t k = 0; k < K; k++) dv[i * K + k] = dvel ? dvel[i * K + k] : mv[i * K + k];
        dspecies[i] = dsp ? dsp[i] : sp[s];
        if (dsp && species_kind(ctx->pot_kind) && (dsp[i] < 0 || (uint32_t)dsp[i] >= nsp)) {
            ctx->err = "species out of range (daughter " + std::to_string(i) + ")";
            return CRM_INVALID_ARGUMENT;
        }
        const bool out_m = reflect_point(P, D, &mp[i * K], &mv[i * K]);
        const bool out_d = reflect_point(P, D, &dp[i * K], &dv[i * K]);
        if (out_m || out_d) {
            ctx->err = "BoundaryError: Particle is out of domain (" + std::string(out_m ? "mother" : "daughter") +
                       " of division " + std::to_string(i) + ")";
            return CRM_BOUNDARY_ERROR;
        }
    }

    // the slot order with the daughters at the end of their mothers' voxels, the
    // voxel each agent moves to, and the identifiers
    std::vector<double> pos2((size_t)K * N2), vel2((size_t)K * N2), hdx2((size_t)2 * K * N2, 0.0),
        hdv2((size_t)2 * K * N2, 0.0);
    std::vector<int32_t> sp2(N2);
    std::vector<uint32_t> id2(N2), vox2(N2), birth2(N2), key2(N2), start2(Vt + 1);
    ctx->pt_ident.resize(3 * N2);
    ctx->pt_parent.resize(3 * N2);
    std::vector<uint64_t> old_ident((size_t)3 * n);
    std::vector<uint64_t> born;
    std::vector<uint32_t> inst_born(ctx->n_inst, 0);
    const uint32_t t_now = (uint32_t)ctx->step;  // a daughter's first push is the next step's
    uint64_t w = 0;
    auto put = [&](uint64_t w, const double* p, const double* v) {
        for (uint32_t k = 0; k < K; k++) {
            pos2[(uint64_t)k * N2 + w] = p[k];
            vel2[(uint64_t)k * N2 + w] = v[k];
        }
    };
    // the voxel a divided agent moves to (its list's voxel g if its position has none)
    auto voxel_key = [&](const InstParams& P, const double* x, uint64_t g) -> uint32_t {
        uint32_t idx[3] = {0, 0, 0};
        bool ok = true;
        for (uint32_t k = 0; k < D; k++) {
            const uint32_t ax = k + 3 - D;
            ok = ok && axis_index_host(x[k], P.min[ax], P.dx[ax], P.n[ax], idx[ax]);
        }
        return ok ? (uint32_t)(P.voxel_offset + ((uint64_t)idx[0] * P.n[1] + idx[1]) * P.n[2] + idx[2])
                  : (uint32_t)g;
    };
    for (uint64_t g = 0; g < Vt; g++) {
        start2[g] = (uint32_t)w;
        born.clear();
        const uint32_t ii = start[g + 1] > start[g] ? instance_of_voxel(ctx, g) : 0u;
        const uint64_t local = g - ctx->hparams[ii].voxel_offset;
        for (uint64_t s = start[g]; s < start[g + 1]; s++, w++) {
            for (uint32_t k = 0; k < K; k++) {
                pos2[(uint64_t)k * N2 + w] = pos[(uint64_t)k * N + s];
                vel2[(uint64_t)k * N2 + w] = vel[(uint64_t)k * N + s];
                for (uint32_t h = 0; h < 2; h++) {
                    hdx2[((uint64_t)h * K + k) * N2 + w] = hdx[((uint64_t)h * K + k) * N + s];
                    hdv2[((uint64_t)h * K + k) * N2 + w] = hdv[((uint64_t)h * K + k) * N + s];
                }
            }
            sp2[w] = sp[s];
            id2[w] = id[s];
            vox2[w] = vox[s];
            birth2[w] = birth[s];
            key2[w] = vox[s];
            if (mother_at[s] < 0) continue;
            const uint64_t i = (uint64_t)mother_at[s], a = id[s];
            put(w, &mp[i * K], &mv[i * K]);
            key2[w] = voxel_key(ctx->hparams[ii], &mp[i * K], g);
            for (int q = 0; q < 3; q++) old_ident[3 * i + q] = ctx->pt_ident[3 * a + q];
            ctx->pt_idc[g] += 1;
            ctx->pt_ident[3 * a] = 1;
            ctx->pt_ident[3 * a + 1] = local;
            ctx->pt_ident[3 * a + 2] = ctx->pt_idc[g];
            for (int q = 0; q < 3; q++) ctx->pt_parent[3 * a + q] = old_ident[3 * i + q];
            born.push_back(i);
        }
        const InstParams& P = ctx->hparams[ii];
        for (const uint64_t i : born) {
            const uint64_t a = N + i;
            put(w, &dp[i * K], &dv[i * K]);
            sp2[w] = dspecies[i];
            id2[w] = (uint32_t)a;
            vox2[w] = (uint32_t)g;
            birth2[w] = t_now;
            key2[w] = voxel_key(P, &dp[i * K], g);
            ctx->pt_idc[g] += 1;
            ctx->pt_ident[3 * a] = 1;
            ctx->pt_ident[3 * a + 1] = local;
            ctx->pt_ident[3 * a + 2] = ctx->pt_idc[g];
            for (int q = 0; q < 3; q++) ctx->pt_parent[3 * a + q] = old_ident[3 * i + q];
            inst_born[ii] += 1;
            w++;
        }
    }
    start2[Vt] = (uint32_t)w;

    // device: room for N2 agents, the new order uploaded, then the re-binning
    if (N2 > ctx->cap_agents || !ctx->pt_births) CRM_TRY(realloc_point_arrays(ctx, N2 + N2 / 8));
    ctx->pt_births = true;
    ctx->N = N2;
    for (uint32_t i = 0; i < ctx->n_inst; i++) ctx->counts[i] += inst_born[i];
    ctx->cur = 0;
    const StateBuf T = ctx->buf[0];
    CRM_HIP(hipMemcpyAsync(T.pos, pos2.data(), pos2.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.vel, vel2.data(), vel2.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.hdx, hdx2.data(), hdx2.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.hdv, hdv2.data(), hdv2.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.species, sp2.data(), N2 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.id, id2.data(), N2 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.vox, vox2.data(), N2 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(T.birth, birth2.data(), N2 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    CRM_HIP(hipMemcpyAsync(ctx->d_start, start2.data(), (Vt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    CRM_HIP(hipMemcpyAsync(ctx->d_key, key2.data(), N2 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    CRM_TRY(rebuild_work(ctx));
    {
        std::vector<ForceTile> tiles;  // the tile shape follows the density
        make_tiles(ctx, ctx->policy, tiles);
        CRM_TRY(upload_tiles(ctx, tiles));
    }
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->sense_parity = -1;
    ctx->graphs_stale = true;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((stay_offset(ctx) + Vt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->d_count, (uint32_t)(stay_offset(ctx) + Vt), 0u);
    CRM_TRY(launch_rank(ctx));
    CRM_TRY(launch_sort(ctx, false, nullptr));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return check_errors(ctx, "agent insertion");
}

}  // namespace

extern "C" {

crm_status crm_grid(const crm_domain* domain, uint32_t dim, uint64_t n_voxels_out[3], double dx_out[3]) {
    if (!domain || !n_voxels_out || !dx_out) return CRM_INVALID_ARGUMENT;
    HostGrid g;
    std::string err;
    crm_status st = host_grid(*domain, dim, g, err);
    if (st) return st;
    for (uint32_t i = 0; i < 3; i++) {
        n_voxels_out[i] = i < dim ? g.n[i + 3 - dim] : 1;
        dx_out[i] = i < dim ? g.dx[i + 3 - dim] : 0.0;
    }
    return CRM_OK;
}

crm_status crm_create(const crm_config* cfg, const crm_instance* instances, uint32_t n_instances, crm_ctx** out) {
    if (!out) return CRM_INVALID_ARGUMENT;
    *out = nullptr;
    g_create_error.clear();
    if (!cfg || !instances || n_instances == 0) {
        g_create_error = "cfg, instances and n_instances > 0 are required";
        return CRM_INVALID_ARGUMENT;
    }
    crm_ctx* ctx = new crm_ctx();
    ctx->cfg = *cfg;
    auto fail = [&](crm_status st, const std::string& msg) {
        g_create_error = msg;
        crm_destroy(ctx);
        return st;
    };
    if (cfg->dim < 2 || cfg->dim > 3) return fail(CRM_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (cfg->solver_order > 2) return fail(CRM_INVALID_ARGUMENT, "mechanics_solver_order must be 0, 1 or 2");
    if (cfg->precision > CRM_PRECISION_MIXED || cfg->weighting > CRM_WEIGHT_PHYSICAL ||
        cfg->history_order > CRM_HISTORY_NEWEST_FIRST || cfg->noise_mode > CRM_NOISE_PHILOX)
        return fail(CRM_INVALID_ARGUMENT, "unknown precision / weighting / history order / noise mode");
    if (!(cfg->dt > 0.0)) return fail(CRM_INVALID_ARGUMENT, "dt must be positive");
    if (cfg->n_subdomains > 65535) return fail(CRM_INVALID_ARGUMENT, "n_subdomains must be below 65536");
    ctx->D = cfg->dim;
    ctx->n_inst = n_instances;
    ctx->inst.assign(instances, instances + n_instances);
    ctx->pot_kind = instances[0].potential.kind;
    ctx->mech_kind = instances[0].mechanics.kind;
    if (ctx->pot_kind > CRM_POTENTIAL_CELL_SORTING_LANGEVIN) return fail(CRM_INVALID_ARGUMENT, "unknown potential kind");
    if (ctx->mech_kind > CRM_MECHANICS_ROD) return fail(CRM_INVALID_ARGUMENT, "unknown mechanics kind");
    ctx->nv = 1;
    if (ctx->mech_kind == CRM_MECHANICS_ROD) {
        ctx->nv = instances[0].mechanics.n_vertices;
        if (ctx->nv < 2 || ctx->nv > CRM_MAX_ROD_VERTICES)
            return fail(CRM_INVALID_ARGUMENT, "rods need 2..16 vertices");
        if (ctx->pot_kind == CRM_POTENTIAL_CELL_SORTING || ctx->pot_kind == CRM_POTENTIAL_CELL_SORTING_LANGEVIN)
            return fail(CRM_INVALID_ARGUMENT, "rods take a point potential (Morse, Mie, BoundLennardJones)");
        ctx->rods_per_block = kBlock / ctx->nv;
        if (cfg->n_subdomains > 1 && cfg->weighting == CRM_WEIGHT_REFERENCE)
            return fail(CRM_INVALID_ARGUMENT, "chili subdomain weighting (n_subdomains > 1) is for point agents");
    }
    ctx->K = ctx->nv * ctx->D;
    uint64_t voff = 0;
    for (uint32_t i = 0; i < n_instances; i++) {
        const crm_instance& I = instances[i];
        if (I.potential.kind != ctx->pot_kind || I.mechanics.kind != ctx->mech_kind ||
            (ctx->mech_kind == CRM_MECHANICS_ROD && I.mechanics.n_vertices != ctx->nv))
            return fail(CRM_INVALID_ARGUMENT, "all instances must share potential and mechanics kinds");
        if (species_kind(I.potential.kind) && (I.potential.n_species == 0 || I.potential.n_species > CRM_MAX_SPECIES))
            return fail(CRM_INVALID_ARGUMENT, "n_species must be 1..4");
        HostGrid g;
        crm_status st = host_grid(I.domain, ctx->D, g, ctx->err);
        if (st) return fail(st, ctx->err);
        ctx->grids.push_back(g);
        InstParams P;
        std::memset(&P, 0, sizeof(P));
        for (int a = 0; a < 3; a++) {
            P.min[a] = g.min[a];
            P.max[a] = g.max[a];
            P.dx[a] = g.dx[a];
            P.dxf[a] = (float)g.dx[a];
            P.qinv[a] = 1073741824.0 / g.dx[a];
            P.inv_dx[a] = 1.0 / g.dx[a];
            P.qunit[a] = (float)(g.dx[a] / 1073741824.0);
            P.n[a] = (uint32_t)g.n[a];
        }
        P.voxel_offset = (uint32_t)voff;
        P.n_voxels = (uint32_t)g.V;
        P.inv_n2 = 1.0 / (double)P.n[2];
        P.inv_n12 = 1.0 / ((double)P.n[1] * (double)P.n[2]);
        if (cfg->n_subdomains > 1 && cfg->weighting == CRM_WEIGHT_REFERENCE) {
            // create_subdomains (cartesian_cuboid_n.rs:493-584): lexicographic chunks
            uint64_t nf = 0, avg = 0;
            if (g.V < cfg->n_subdomains || !decomp_res(g.V, cfg->n_subdomains, nf, avg))
                return fail(CRM_SETUP_ERROR, "DecomposeError: Could not find a suiting decomposition of " +
                                                 std::to_string(g.V) + " voxels into " +
                                                 std::to_string(cfg->n_subdomains) + " subdomains");
            P.sub_on = 1;
            P.sub_avg = (uint32_t)avg;
            P.sub_avg1 = (uint32_t)std::max<uint64_t>(avg - 1, 1);
            P.sub_switch = (uint32_t)(nf * avg);
            P.sub_n = (uint32_t)nf;
        }
        fill_pot<double>(P.pd, I.potential);
        fill_pot<float>(P.pf, I.potential);
        const crm_mechanics& M = I.mechanics;
        P.mass = M.mass;
        P.damping = M.damping;
        P.kb_temperature = M.kb_temperature;
        P.diffusion_constant = M.diffusion_constant;
        P.spring_tension = M.spring_tension;
        P.rigidity = M.rigidity;
        P.spring_length = M.spring_length;
        P.inv_mass = 1.0 / M.mass;
        P.langevin_coeff = std::sqrt(2.0 * M.damping * M.kb_temperature / M.mass);
        P.brownian_q = M.diffusion_constant / M.kb_temperature;
        P.brownian_coeff = std::sqrt(2.0 * M.diffusion_constant);
        ctx->hparams.push_back(P);
        voff += g.V;
        if (I.potential.kind == CRM_POTENTIAL_MORSE_SPECIES || I.potential.kind == CRM_POTENTIAL_CELL_SORTING_LANGEVIN) {
            for (uint32_t a = 0; a < I.potential.n_species; a++)
                for (uint32_t b = 0; b < I.potential.n_species; b++)
                    if (I.potential.species_matrix[a * CRM_MAX_SPECIES + b] !=
                        I.potential.species_matrix[b * CRM_MAX_SPECIES + a])
                        ctx->symmetric = false;
        }
    }
    if (voff >= 0xffffffffull) return fail(CRM_INVALID_ARGUMENT, "too many voxels (more than 2^32 - 1)");
    ctx->Vt = voff;
    ctx->device = cfg->device;
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return fail(CRM_DEVICE_ERROR, std::string("hipSetDevice: ") + hipGetErrorString(e));
    e = scratch_address(cfg->device, ctx->d_noalloc);
    if (e != hipSuccess) return fail(CRM_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ctx->inst_host.assign(n_instances, 0);
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(CRM_DEVICE_ERROR, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    if (cfg->flags & CRM_FLAG_TIMING) {
        for (int i = 0; i < crm_ctx::kTimingPairs; i++)
            if (hipEventCreate(&ctx->ev_a[i]) != hipSuccess || hipEventCreate(&ctx->ev_b[i]) != hipSuccess)
                return fail(CRM_DEVICE_ERROR, "hipEventCreate failed");
    }
    if (hipMalloc((void**)&ctx->d_ip, sizeof(InstParams) * n_instances) != hipSuccess ||
        hipMalloc((void**)&ctx->d_ctrl, sizeof(Control)) != hipSuccess)
        return fail(CRM_DEVICE_ERROR, "hipMalloc failed");
    e = hipMemcpy(ctx->d_ip, ctx->hparams.data(), sizeof(InstParams) * n_instances, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(CRM_DEVICE_ERROR, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *out = ctx;
    return CRM_OK;
}

crm_status crm_set_agents(crm_ctx* ctx, const uint64_t* counts, const double* pos, const double* vel,
                          const int32_t* species) {
    if (!ctx || !counts || !pos) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    uint64_t N = 0;
    for (uint32_t i = 0; i < ctx->n_inst; i++) N += counts[i];
    if (N == 0) {
        ctx->err = "at least one agent is required";
        return CRM_INVALID_ARGUMENT;
    }
    if (N >= 0xffffffffull) {
        ctx->err = "too many agents (more than 2^32 - 1)";
        return CRM_INVALID_ARGUMENT;
    }
    const uint32_t nsp = ctx->inst[0].potential.n_species;
    if (species && species_kind(ctx->pot_kind))
        for (uint64_t a = 0; a < N; a++)
            if (species[a] < 0 || (uint32_t)species[a] >= nsp) {
                ctx->err = "species out of range (agent " + std::to_string(a) + ")";
                return CRM_INVALID_ARGUMENT;
            }
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    free_device(ctx);
    ctx->N = N;
    ctx->counts.assign(counts, counts + ctx->n_inst);
    ctx->have_agents = false;
    ctx->halted = false;
    ctx->step = 0;
    ctx->err.clear();
    const uint32_t K = ctx->K;

    // block work table
    std::vector<BlockWork> work;
    uint64_t first = 0;
    for (uint32_t i = 0; i < ctx->n_inst; i++) {
        ctx->hparams[i].agent_offset = first;
        ctx->hparams[i].n_agents = counts[i];
        const uint64_t per = is_rod(ctx) ? ctx->rods_per_block : (uint64_t)kBlock;
        for (uint64_t b = 0; b < counts[i]; b += per) {
            BlockWork w;
            w.inst = i;
            w.first = first + b;
            w.count = (uint32_t)std::min<uint64_t>(per, counts[i] - b);
            work.push_back(w);
        }
        first += counts[i];
    }
    ctx->n_work = (uint32_t)work.size();
    std::vector<ForceTile> tiles;
    make_tiles(ctx, ctx->policy, tiles);
    CRM_HIP(hipMemcpy(ctx->d_ip, ctx->hparams.data(), sizeof(InstParams) * ctx->n_inst, hipMemcpyHostToDevice));

    const uint64_t Vt = ctx->Vt;
    ctx->n_tiles = (uint32_t)((Vt + kScanTile - 1) / kScanTile);
    // room for the work blocks of `C` agents (a growing cycle rebuilds the table)
    {
        const uint64_t per = is_rod(ctx) ? ctx->rods_per_block : (uint64_t)kBlock;
        const uint64_t Cw = (is_rod(ctx) && ctx->cycle_on) ? std::max<uint64_t>(N, ctx->cyc.max_agents) : N;
        CRM_TRY(dalloc(ctx, ctx->d_work, std::max<size_t>(work.size(), Cw / per + ctx->n_inst + 1)));
    }
    CRM_TRY(dalloc(ctx, ctx->d_start, Vt + 1));
    CRM_TRY(dalloc(ctx, ctx->d_count, stay_offset(ctx) + Vt));  // arrivals, then stayers
    CRM_TRY(dalloc(ctx, ctx->d_start_local, Vt));
    CRM_TRY(dalloc(ctx, ctx->d_bsum, ctx->n_tiles));
    // N-sized arrays hold `C` agents: the rods of a cycle grow up to max_agents
    const uint64_t C = (is_rod(ctx) && ctx->cycle_on) ? std::max<uint64_t>(N, ctx->cyc.max_agents) : N;
    ctx->cap_agents = C;
    CRM_TRY(dalloc(ctx, ctx->d_key, 2 * C));
    CRM_TRY(dalloc(ctx, ctx->d_rank, C));
    CRM_TRY(dalloc(ctx, ctx->d_perm, C));
    CRM_TRY(dalloc(ctx, ctx->d_force, (size_t)K * C));
    if (is_rod(ctx)) {
        CRM_TRY(dalloc(ctx, ctx->d_recrod, (size_t)K * C));
        CRM_TRY(dalloc(ctx, ctx->d_bnd, C));
        CRM_TRY(dalloc(ctx, ctx->d_seg, (size_t)(ctx->nv - 1) * C));
    }
    else if (ctx->cfg.precision == CRM_PRECISION_MIXED) CRM_TRY(dalloc(ctx, ctx->d_recq, N));
    else CRM_TRY(dalloc(ctx, ctx->d_rec64, N));
    if (ctx->sense_on) CRM_TRY(dalloc(ctx, ctx->d_sense, (size_t)ctx->nv * C));
    for (auto& b : ctx->buf) {
        CRM_TRY(dalloc(ctx, b.pos, (size_t)K * C));
        CRM_TRY(dalloc(ctx, b.vel, (size_t)K * C));
        CRM_TRY(dalloc(ctx, b.hdx, (size_t)2 * K * C));
        CRM_TRY(dalloc(ctx, b.hdv, (size_t)2 * K * C));
        CRM_TRY(dalloc(ctx, b.species, C));
        CRM_TRY(dalloc(ctx, b.id, C));
        CRM_TRY(dalloc(ctx, b.vox, C));
        if (is_rod(ctx)) {
            CRM_TRY(dalloc(ctx, b.spring, C));
            CRM_TRY(dalloc(ctx, b.birth, C));
            CRM_TRY(dalloc(ctx, b.growth, C));
            CRM_TRY(dalloc(ctx, b.ident, 3 * C));
            CRM_TRY(dalloc(ctx, b.parent, 3 * C));
        }
    }
    if (is_rod(ctx) && ctx->cycle_on) {
        CRM_TRY(dalloc(ctx, ctx->d_divflag, C));
        CRM_TRY(dalloc(ctx, ctx->d_drank, C));
        CRM_TRY(dalloc(ctx, ctx->d_dctr, C));
        CRM_TRY(dalloc(ctx, ctx->d_ndiv, Vt + 1));
        CRM_TRY(dalloc(ctx, ctx->d_dbefore, Vt + 1));
        CRM_TRY(dalloc(ctx, ctx->d_vcount, Vt));
        CRM_TRY(dalloc(ctx, ctx->d_inst_div, ctx->n_inst));
        CRM_TRY(alloc_scan_tmp(ctx));
    }
    CRM_HIP(hipMemcpy(ctx->d_work, work.data(), sizeof(BlockWork) * work.size(), hipMemcpyHostToDevice));
    ctx->work_host = work;  // later rebuilds (removals) shrink it in place
    CRM_TRY(upload_tiles(ctx, tiles));
    CRM_HIP(hipMemsetAsync(ctx->d_count, 0, sizeof(uint32_t) * (stay_offset(ctx) + Vt), ctx->stream));

    // upload in agent-index order (slot = agent index before the first sort)
    ctx->cur = 0;
    StateBuf& S = ctx->buf[0];
    {
        std::vector<double> tp((size_t)K * N), tv((size_t)K * N, 0.0);
        for (uint64_t a = 0; a < N; a++)
            for (uint32_t k = 0; k < K; k++) {
                tp[(uint64_t)k * N + a] = pos[a * K + k];
                if (vel) tv[(uint64_t)k * N + a] = vel[a * K + k];
            }
        std::vector<int32_t> sp(N, 0);
        std::vector<uint32_t> id(N);
        for (uint64_t a = 0; a < N; a++) {
            sp[a] = species ? species[a] : 0;
            id[a] = (uint32_t)a;
        }
        CRM_HIP(hipMemcpyAsync(S.pos, tp.data(), tp.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CRM_HIP(hipMemcpyAsync(S.vel, tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CRM_HIP(hipMemcpyAsync(S.species, sp.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        CRM_HIP(hipMemcpyAsync(S.id, id.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        if (is_rod(ctx)) {
            // per-rod spring length from the instance, first push at step 0,
            // growth at the cycle's base rate, CellIdentifier::Initial(index in
            // the instance), no parent
            std::vector<double> sl(N), gr(N, ctx->cycle_on ? ctx->cyc.growth_base_rate : 0.0);
            std::vector<uint32_t> birth(N, 0);
            std::vector<uint64_t> ident(3 * N, 0), parent(3 * N, 0);
            uint64_t a = 0;
            for (uint32_t i = 0; i < ctx->n_inst; i++)
                for (uint64_t j = 0; j < counts[i]; j++, a++) {
                    sl[a] = ctx->hparams[i].spring_length;
                    ident[N + a] = j;
                    parent[a] = 3;
                }
            CRM_HIP(hipMemcpyAsync(S.spring, sl.data(), N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            CRM_HIP(hipMemcpyAsync(S.growth, gr.data(), N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            CRM_HIP(hipMemcpyAsync(S.birth, birth.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
            CRM_HIP(hipMemcpyAsync(S.ident, ident.data(), 3 * N * sizeof(uint64_t), hipMemcpyHostToDevice,
                                   ctx->stream));
            CRM_HIP(hipMemcpyAsync(S.parent, parent.data(), 3 * N * sizeof(uint64_t), hipMemcpyHostToDevice,
                                   ctx->stream));
        }
        CRM_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->pt_ident.clear();
    ctx->pt_parent.clear();
    if (!is_rod(ctx)) {  // CellIdentifier::Initial(index among the instance's cells), no parent
        ctx->pt_ident.assign(3 * N, 0);
        ctx->pt_parent.assign(3 * N, 0);
        uint64_t a = 0;
        for (uint32_t i = 0; i < ctx->n_inst; i++)
            for (uint64_t j = 0; j < counts[i]; j++, a++) {
                ctx->pt_ident[3 * a + 1] = j;
                ctx->pt_parent[3 * a] = 3;
            }
    }
    CRM_TRY(reset_control(ctx, 0));
    Arrays A = arrays(ctx);
    if (ctx->D == 2)
        hipLaunchKernelGGL(k_initial_key<2>, dim3(ctx->n_work), dim3(kBlock), 0, ctx->stream, A, S, ctx->nv);
    else
        hipLaunchKernelGGL(k_initial_key<3>, dim3(ctx->n_work), dim3(kBlock), 0, ctx->stream, A, S, ctx->nv);
    CRM_HIP(hipGetLastError());
    CRM_TRY(check_errors(ctx, "initial placement"));
    CRM_TRY(launch_sort(ctx, false, nullptr));
    if (is_rod(ctx) && ctx->cycle_on)
        hipLaunchKernelGGL(k_init_vcount, dim3((uint32_t)((Vt + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                           ctx->d_start, ctx->d_vcount, (uint32_t)Vt);
    ctx->pt_births = false;
    ctx->pt_idc.clear();
    if (!is_rod(ctx)) {  // Voxel::id_counter after insert_initial_cells (datastructures.rs:286-297)
        std::vector<uint32_t> st(Vt + 1);
        CRM_HIP(hipMemcpyAsync(st.data(), ctx->d_start, (Vt + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        ctx->pt_idc.resize(Vt);
        for (uint64_t v = 0; v < Vt; v++) ctx->pt_idc[v] = st[v + 1] - st[v];
    }
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_agents = true;
    return CRM_OK;
}

crm_status crm_step(crm_ctx* ctx, uint32_t n_steps, const double* noise) {
    if (!ctx) return CRM_INVALID_ARGUMENT;
    if (!ctx->have_agents) {
        ctx->err = "crm_set_agents must be called first";
        return CRM_INVALID_ARGUMENT;
    }
    if (ctx->halted) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    const bool use_noise_buf = ctx->cfg.noise_mode == CRM_NOISE_BUFFER && noise != nullptr;
    const size_t per_step = (size_t)ctx->N * ctx->K;
    if (use_noise_buf) {
        size_t need = per_step * n_steps;
        if (need > ctx->noise_cap) {
            if (ctx->d_noise) CRM_HIP(hipFree(ctx->d_noise));
            ctx->d_noise = nullptr;
            CRM_TRY(dalloc(ctx, ctx->d_noise, need));
            ctx->noise_cap = need;
        }
        CRM_HIP(hipMemcpyAsync(ctx->d_noise, noise, need * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    if (ctx->cycle_on) {  // the grown order is sized on the host every step
        for (uint32_t k = 0; k < n_steps; k++) {
            // errors of a step surface at the next step's control read or below
            crm_status st = launch_cycle_step(ctx, use_noise_buf ? ctx->d_noise + per_step * k : nullptr);
            if (st) return st;
            ctx->step += 1;
            ctx->sense_parity = -1;  // counts are by the pre-step order
        }
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        return check_errors(ctx, "step");
    }
    const bool graphs = (ctx->cfg.flags & CRM_FLAG_GRAPH) && !timing_on(ctx) && !use_noise_buf;
    if (ctx->graphs_stale) {  // captured before a removal changed the agent count
        for (auto& g : ctx->graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        ctx->graphs_stale = false;
    }
    uint32_t s = 0;
    if (graphs) {  // each graph launch costs the device a few microseconds
        while (s + kLongGraphSteps <= n_steps) {
            if (!ctx->graph[2 + ctx->cur]) CRM_TRY(build_graph(ctx, ctx->cur, true));
            CRM_HIP(hipGraphLaunch(ctx->graph[2 + ctx->cur], ctx->stream));
            s += kLongGraphSteps;
        }
        while (s + 2 <= n_steps) {
            if (!ctx->graph[ctx->cur]) CRM_TRY(build_graph(ctx, ctx->cur, false));
            CRM_HIP(hipGraphLaunch(ctx->graph[ctx->cur], ctx->stream));
            s += 2;
        }
    }
    for (; s < n_steps; s++) CRM_TRY(launch_step(ctx, use_noise_buf ? ctx->d_noise + per_step * s : nullptr));
    Control c;
    CRM_TRY(read_control(ctx, c));
    // The steps after a failing one were launched but did nothing: the state is
    // the failing step's end, in the buffer that step wrote.
    uint64_t done = n_steps;
    if (c.error_flags) {
        const uint64_t t_err = c.error_key >> 32;
        if (t_err >= ctx->step && t_err < ctx->step + n_steps) {
            done = t_err + 1 - ctx->step;
            if ((n_steps - done) & 1) ctx->cur = 1 - ctx->cur;
        }
    }
    ctx->step += done;
    // the last step's forces (and counts) read the buffer before its gather
    if (done > 0) ctx->sense_parity = 1 - ctx->cur;
    return check_errors(ctx, "step");
}

crm_status crm_compute_forces(crm_ctx* ctx, double* force) {
    if (!ctx || !force) return CRM_INVALID_ARGUMENT;
    if (!ctx->have_agents) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_TRY(launch_force(ctx));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    crm_status st = check_errors(ctx, "forces");
    if (st) return st;
    std::vector<uint32_t> id;
    CRM_TRY(download_ids(ctx, id));
    if (!force_f32(ctx)) return download_by_id(ctx, ctx->d_force, ctx->K, force, id);
    std::vector<float> tmp((size_t)ctx->K * ctx->N);
    CRM_HIP(hipMemcpyAsync(tmp.data(), ctx->d_force, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t s = 0; s < ctx->N; s++)
        for (uint32_t k = 0; k < ctx->K; k++) force[(uint64_t)id[s] * ctx->K + k] = (double)tmp[(uint64_t)k * ctx->N + s];
    return CRM_OK;
}

crm_status crm_get_state(crm_ctx* ctx, double* pos, double* vel) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> id;
    CRM_TRY(download_ids(ctx, id));
    if (pos) CRM_TRY(download_by_id(ctx, ctx->buf[ctx->cur].pos, ctx->K, pos, id));
    if (vel) CRM_TRY(download_by_id(ctx, ctx->buf[ctx->cur].vel, ctx->K, vel, id));
    return CRM_OK;
}

crm_status crm_get_history(crm_ctx* ctx, double* dx_hist, double* dv_hist, uint32_t* n_stored) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    const uint32_t cap = ctx->cfg.solver_order;
    const uint64_t t = ctx->step;
    const uint32_t size = cap == 0 ? 0u : (uint32_t)std::min<uint64_t>(t, cap);
    const uint32_t oldest = (size < cap || cap == 0) ? 0u : (uint32_t)(t % cap);
    if (n_stored) *n_stored = size;
    std::vector<uint32_t> id;
    CRM_TRY(download_ids(ctx, id));
    const uint64_t slot_elems = (uint64_t)ctx->N * ctx->K;
    if (ctx->pt_births) {  // inserted agents hold the pushes since their birth, oldest first, then zeros
        const uint64_t N = ctx->N;
        const uint32_t K = ctx->K;
        std::vector<double> hx(2 * slot_elems), hv(2 * slot_elems);
        std::vector<uint32_t> birth(N);
        const StateBuf& S = ctx->buf[ctx->cur];
        CRM_HIP(hipMemcpyAsync(hx.data(), S.hdx, hx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CRM_HIP(hipMemcpyAsync(hv.data(), S.hdv, hv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CRM_HIP(hipMemcpyAsync(birth.data(), S.birth, N * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t s = 0; s < N; s++) {
            const uint64_t b = std::min<uint64_t>(birth[s], t);
            const uint32_t sz = cap == 0 ? 0u : (uint32_t)std::min<uint64_t>(t - b, cap);
            for (uint32_t i = 0; i < 2; i++)
                for (uint32_t k = 0; k < K; k++) {
                    const uint64_t o = i * slot_elems + (uint64_t)id[s] * K + k;
                    double a = 0.0, c = 0.0;
                    if (i < sz) {
                        const uint64_t slot = (t - sz + i) % cap;
                        a = hx[(slot * K + k) * N + s];
                        c = hv[(slot * K + k) * N + s];
                    }
                    if (dx_hist) dx_hist[o] = a;
                    if (dv_hist) dv_hist[o] = c;
                }
        }
        return CRM_OK;
    }
    for (uint32_t i = 0; i < 2; i++) {
        if (i < size) {
            uint32_t slot = (oldest + i) % cap;
            if (dx_hist)
                CRM_TRY(download_by_id(ctx, ctx->buf[ctx->cur].hdx + slot * slot_elems, ctx->K,
                                       dx_hist + i * slot_elems, id));
            if (dv_hist)
                CRM_TRY(download_by_id(ctx, ctx->buf[ctx->cur].hdv + slot * slot_elems, ctx->K,
                                       dv_hist + i * slot_elems, id));
        } else {
            if (dx_hist) std::fill(dx_hist + i * slot_elems, dx_hist + (i + 1) * slot_elems, 0.0);
            if (dv_hist) std::fill(dv_hist + i * slot_elems, dv_hist + (i + 1) * slot_elems, 0.0);
        }
    }
    return CRM_OK;
}

crm_status crm_get_voxels(crm_ctx* ctx, uint32_t* voxel_of_agent, uint32_t* voxel_start) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    if (voxel_of_agent) {
        std::vector<uint32_t> id, vox(ctx->N);
        CRM_TRY(download_ids(ctx, id));
        CRM_HIP(hipMemcpyAsync(vox.data(), ctx->buf[ctx->cur].vox, ctx->N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t s = 0; s < ctx->N; s++) voxel_of_agent[id[s]] = vox[s];
    }
    if (voxel_start) {
        CRM_HIP(hipMemcpyAsync(voxel_start, ctx->d_start, (ctx->Vt + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
    }
    return CRM_OK;
}

crm_status crm_get_voxel_order(crm_ctx* ctx, uint32_t* order) {
    if (!ctx || !order || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_HIP(hipMemcpyAsync(order, ctx->buf[ctx->cur].id, ctx->N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

}  // extern "C"

// ---- tiling autotuner ------------------------------------------------------
namespace {

std::mutex g_tune_mutex;
std::map<std::string, uint32_t> g_tune_cache;  // workload key -> tiling policy

std::string tune_key(const crm_ctx* ctx) {
    char buf[256];
    const InstParams& P = ctx->hparams[0];
    std::snprintf(buf, sizeof buf, "dim=%u precision=%u potential=%u weighting=%u instances=%u agents=%llu grid=%ux%ux%u",
                  ctx->D, ctx->cfg.precision, ctx->pot_kind, ctx->cfg.weighting, ctx->n_inst,
                  (unsigned long long)ctx->N, P.n[0], P.n[1], P.n[2]);
    return buf;
}

// The cache file is a flat JSON object {"<key>": <policy>, ...}.
void read_tune_file(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return;
    std::string text;
    char chunk[4096];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) text.append(chunk, n);
    std::fclose(f);
    size_t i = 0;
    while ((i = text.find('"', i)) != std::string::npos) {
        const size_t j = text.find('"', i + 1);
        if (j == std::string::npos) break;
        const std::string key = text.substr(i + 1, j - i - 1);
        const size_t c = text.find(':', j);
        if (c == std::string::npos) break;
        char* end = nullptr;
        const unsigned long v = std::strtoul(text.c_str() + c + 1, &end, 10);
        if (end != text.c_str() + c + 1 && v < kNumTilePolicies) g_tune_cache[key] = (uint32_t)v;
        i = (size_t)(end - text.c_str());
    }
}

void write_tune_file(const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return;
    std::fputs("{\n", f);
    size_t k = 0;
    for (const auto& e : g_tune_cache)
        std::fprintf(f, "  \"%s\": %u%s\n", e.first.c_str(), e.second, ++k < g_tune_cache.size() ? "," : "");
    std::fputs("}\n", f);
    std::fclose(f);
}

crm_status apply_policy(crm_ctx* ctx, uint32_t pol) {
    std::vector<ForceTile> tiles;
    make_tiles(ctx, pol, tiles);
    CRM_TRY(upload_tiles(ctx, tiles));
    ctx->policy = pol;
    for (auto& g : ctx->graph)  // captured with the old tile table
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return CRM_OK;
}

}  // namespace

extern "C" {

crm_status crm_autotune(crm_ctx* ctx, const char* cache_path, uint32_t* chosen) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    if (ctx->halted) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    if (is_rod(ctx)) {  // rods do not use force tiles
        if (chosen) *chosen = ctx->policy;
        return CRM_OK;
    }
    const std::string key = tune_key(ctx);
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        if (cache_path) read_tune_file(cache_path);
        const auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            CRM_TRY(apply_policy(ctx, it->second));
            if (chosen) *chosen = it->second;
            return CRM_OK;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CRM_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        ctx->err = "hipEventCreate failed";
        return CRM_DEVICE_ERROR;
    }
    uint32_t best = ctx->policy;
    float best_ms = 0.0f;
    crm_status st = CRM_OK;
    for (uint32_t p = 0; p < kNumTilePolicies && st == CRM_OK; p++) {
        st = apply_policy(ctx, p);
        if (st) break;
        st = launch_force(ctx);  // warm-up
        if (st) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) st = CRM_DEVICE_ERROR;
        for (int r = 0; r < 3 && st == CRM_OK; r++) st = launch_force(ctx);
        if (st) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = CRM_DEVICE_ERROR;
        if (st == CRM_OK && (p == 0 || ms < best_ms)) {
            best_ms = ms;
            best = p;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st == CRM_DEVICE_ERROR && ctx->err.empty()) ctx->err = "HIP error while timing the tiling policies";
    if (st) return st;
    CRM_TRY(apply_policy(ctx, best));
    CRM_TRY(check_errors(ctx, "autotune"));
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune_cache[key] = best;
        if (cache_path) write_tune_file(cache_path);
    }
    if (chosen) *chosen = best;
    return CRM_OK;
}

crm_status crm_set_domain_force(crm_ctx* ctx, uint32_t inst, const crm_domain_force* f) {
    if (!ctx || !f || inst >= ctx->n_inst) return CRM_INVALID_ARGUMENT;
    if (f->kind > CRM_DOMAIN_FORCE_ROD_GEL) {
        ctx->err = "unknown domain force kind";
        return CRM_INVALID_ARGUMENT;
    }
    if (f->kind == CRM_DOMAIN_FORCE_ROD_GEL && !(is_rod(ctx) && ctx->D == 3)) {
        ctx->err = "the rod gel force (CartesianSubDomainRods) acts on 3-D rods only";
        return CRM_INVALID_ARGUMENT;
    }
    InstParams& P = ctx->hparams[inst];
    P.dforce = f->kind;
    for (int k = 0; k < 3; k++) {
        P.dk[k] = f->strength[k];
        P.dc[k] = f->center[k];
    }
    P.gel = f->gel_pressure;
    P.fric = f->surface_friction;
    P.fric_d = f->surface_friction_distance;
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    CRM_HIP(hipMemcpy(ctx->d_ip + inst, &P, sizeof(InstParams), hipMemcpyHostToDevice));
    return CRM_OK;
}

crm_status crm_set_rod_cycle(crm_ctx* ctx, const crm_rod_cycle* y) {
    if (!ctx || !y) return CRM_INVALID_ARGUMENT;
    if (!is_rod(ctx)) {
        ctx->err = "the rods example's cycle needs rod agents";
        return CRM_INVALID_ARGUMENT;
    }
    if (ctx->have_agents) {
        ctx->err = "crm_set_rod_cycle must come before crm_set_agents (it sizes the context for max_agents rods)";
        return CRM_INVALID_ARGUMENT;
    }
    const bool ok = std::isfinite(y->growth_base_rate) && std::isfinite(y->division_threshold) &&
                    std::isfinite(y->division_radius) && std::isfinite(y->sensing_radius) &&
                    y->neighbor_saturation > 0.0 && std::isfinite(y->neighbor_saturation) && y->max_agents > 0 &&
                    y->max_agents < 0xffffffffull;
    if (!ok) {
        ctx->err = "invalid cycle parameters";
        return CRM_INVALID_ARGUMENT;
    }
    CRM_TRY(crm_set_neighbor_sensing(ctx, 1, y->sensing_radius));
    ctx->cyc = *y;
    ctx->cycle_on = true;
    return CRM_OK;
}

crm_status crm_set_rod_spring_lengths(crm_ctx* ctx, const double* sl) {
    if (!ctx || !sl || !ctx->have_agents || !is_rod(ctx)) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> id;
    CRM_TRY(download_ids(ctx, id));
    std::vector<double> v(ctx->N);
    for (uint64_t s = 0; s < ctx->N; s++) v[s] = sl[id[s]];
    CRM_HIP(hipMemcpyAsync(ctx->buf[ctx->cur].spring, v.data(), ctx->N * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

crm_status crm_get_cycle_state(crm_ctx* ctx, double* spring_length, double* growth_rate, uint64_t* identifier,
                               uint64_t* parent, uint32_t* history_held) {
    if (!ctx || !ctx->have_agents || !is_rod(ctx)) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> id;
    CRM_TRY(download_ids(ctx, id));
    const uint64_t N = ctx->N;
    const StateBuf& S = ctx->buf[ctx->cur];
    if (spring_length) CRM_TRY(download_by_id(ctx, S.spring, 1, spring_length, id));
    if (growth_rate) CRM_TRY(download_by_id(ctx, S.growth, 1, growth_rate, id));
    std::vector<uint64_t> tmp(3 * N);
    uint64_t* outs[2] = {identifier, parent};
    uint64_t* srcs[2] = {S.ident, S.parent};
    for (int q = 0; q < 2; q++) {
        if (!outs[q]) continue;
        CRM_HIP(hipMemcpyAsync(tmp.data(), srcs[q], 3 * N * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t s = 0; s < N; s++)
            for (int j = 0; j < 3; j++) outs[q][(uint64_t)id[s] * 3 + j] = tmp[(uint64_t)j * N + s];
    }
    if (history_held) {
        std::vector<uint32_t> birth(N);
        CRM_HIP(hipMemcpyAsync(birth.data(), S.birth, N * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t cap = ctx->cfg.solver_order;
        for (uint64_t s = 0; s < N; s++) {
            const uint64_t pushes = ctx->step > birth[s] ? ctx->step - birth[s] : 0;
            history_held[id[s]] = (uint32_t)std::min(pushes, cap);
        }
    }
    return CRM_OK;
}

crm_status crm_get_identifiers(crm_ctx* ctx, uint64_t* identifier, uint64_t* parent, uint32_t* history_held) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    if (is_rod(ctx)) return crm_get_cycle_state(ctx, nullptr, nullptr, identifier, parent, history_held);
    const uint64_t N = ctx->N;
    if (identifier) std::copy(ctx->pt_ident.begin(), ctx->pt_ident.end(), identifier);
    if (parent) std::copy(ctx->pt_parent.begin(), ctx->pt_parent.end(), parent);
    if (history_held) {
        const uint32_t cap = ctx->cfg.solver_order;
        const uint32_t held = (uint32_t)std::min<uint64_t>(ctx->step, cap);
        std::fill(history_held, history_held + N, held);
        if (ctx->pt_births) {  // daughters hold the pushes since their birth
            CRM_HIP(hipSetDevice(ctx->device));
            std::vector<uint32_t> id(N), birth(N);
            CRM_HIP(hipMemcpyAsync(id.data(), ctx->buf[ctx->cur].id, N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
            CRM_HIP(hipMemcpyAsync(birth.data(), ctx->buf[ctx->cur].birth, N * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
            CRM_HIP(hipStreamSynchronize(ctx->stream));
            for (uint64_t s = 0; s < N; s++)
                history_held[id[s]] = (uint32_t)std::min<uint64_t>(ctx->step - std::min<uint64_t>(birth[s], ctx->step), cap);
        }
    }
    return CRM_OK;
}

crm_status crm_remove_agents(crm_ctx* ctx, const uint8_t* remove) {
    uint64_t n_rm = 0;
    CRM_TRY(remove_prelude(ctx, remove, n_rm));
    if (n_rm == 0) return CRM_OK;
    RemoveLayout L;
    CRM_TRY(remove_layout(ctx, ctx->N, L));
    if (L.total > ctx->rm_cap) {  // grow-only scratch of the context
        if (ctx->d_rm) CRM_HIP(hipFree(ctx->d_rm));
        ctx->d_rm = nullptr;
        ctx->rm_cap = 0;
        CRM_TRY(dalloc(ctx, ctx->d_rm, L.total));
        ctx->rm_cap = L.total;
    }
    return remove_agents_impl(ctx, remove, n_rm, ctx->d_rm, L);
}

crm_status crm_remove_agents_noalloc(crm_ctx* ctx, const uint8_t* remove) {
    uint64_t n_rm = 0;
    CRM_TRY(remove_prelude(ctx, remove, n_rm));
    if (n_rm == 0) return CRM_OK;
    if (ctx->N > kNoAllocMaxAgents || ctx->n_inst > kNoAllocMaxInstances) {
        ctx->err = "crm_remove_agents_noalloc handles at most " + std::to_string(kNoAllocMaxAgents) +
                   " agents in " + std::to_string(kNoAllocMaxInstances) + " instances (the context holds " +
                   std::to_string(ctx->N) + " in " + std::to_string(ctx->n_inst) + ")";
        return CRM_INVALID_ARGUMENT;
    }
    RemoveLayout L;
    CRM_TRY(remove_layout(ctx, ctx->N, L));
    if (L.total > kNoAllocScratchBytes || L.scan_bytes > kNoAllocScanReserve) {
        ctx->err = "crm_remove_agents_noalloc: the scan needs more scratch than reserved";
        return CRM_INVALID_ARGUMENT;
    }
    return remove_agents_impl(ctx, remove, n_rm, ctx->d_noalloc, L);
}

crm_status crm_insert_agents(crm_ctx* ctx, uint64_t n, const uint64_t* mother, const double* mother_pos,
                             const double* mother_vel, const double* daughter_pos, const double* daughter_vel,
                             const int32_t* daughter_species) {
    if (!ctx) return CRM_INVALID_ARGUMENT;
    if (!ctx->have_agents) {
        ctx->err = "crm_set_agents must be called first";
        return CRM_INVALID_ARGUMENT;
    }
    if (ctx->halted) return CRM_INVALID_ARGUMENT;
    if (is_rod(ctx)) {
        ctx->err = "rods divide through the rods example's cycle (crm_set_rod_cycle)";
        return CRM_INVALID_ARGUMENT;
    }
    if (n == 0) return CRM_OK;
    if (!mother || !daughter_pos) return CRM_INVALID_ARGUMENT;
    crm_status st = insert_agents_impl(ctx, n, mother, mother_pos, mother_vel, daughter_pos, daughter_vel,
                                       daughter_species);
    if (st == CRM_DEVICE_ERROR) ctx->halted = true;
    return st;
}

crm_status crm_set_neighbor_sensing(crm_ctx* ctx, uint32_t enable, double radius) {
    if (!ctx) return CRM_INVALID_ARGUMENT;
    if (enable && std::isnan(radius)) {
        ctx->err = "the sensing radius is NaN";
        return CRM_INVALID_ARGUMENT;
    }
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    if (enable && radius > 0.0) {
        // smallest t with sqrt(t) >= radius: then sqrt(d2) < radius <=> d2 < t
        // (sqrt is correctly rounded and monotone)
        t = radius * radius;
        while (std::sqrt(t) < radius) t = std::nextafter(t, HUGE_VAL);
        for (double p = std::nextafter(t, 0.0); t > 0.0 && std::sqrt(p) >= radius; p = std::nextafter(t, 0.0))
            t = p;
    }
    if (enable && ctx->have_agents && !ctx->d_sense) CRM_TRY(dalloc(ctx, ctx->d_sense, (size_t)ctx->nv * ctx->N));
    ctx->sense_on = enable != 0;
    ctx->sense_t = t;
    ctx->sense_parity = -1;
    for (auto& g : ctx->graph)  // captured with the old kernel arguments
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return CRM_OK;
}

crm_status crm_get_neighbor_counts(crm_ctx* ctx, uint32_t* counts) {
    if (!ctx || !counts || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    if (!ctx->sense_on || ctx->sense_parity < 0) {
        ctx->err = ctx->cycle_on && ctx->step > 0
                       ? "with the rods cycle the counts are consumed by the step (the growth rates of "
                         "crm_get_cycle_state); crm_compute_forces evaluates them for the current state"
                       : "no force evaluation with neighbour sensing since it was enabled or the agents were set";
        return CRM_INVALID_ARGUMENT;
    }
    CRM_HIP(hipSetDevice(ctx->device));
    const uint64_t N = ctx->N, nv = ctx->nv;
    std::vector<uint32_t> id(N), c((size_t)nv * N);
    CRM_HIP(hipMemcpyAsync(id.data(), ctx->buf[ctx->sense_parity].id, N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    CRM_HIP(hipMemcpyAsync(c.data(), ctx->d_sense, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t s = 0; s < N; s++) {
        uint32_t n = 0;
        for (uint64_t v = 0; v < nv; v++) n += c[s * nv + v];
        counts[id[s]] = n;
    }
    return CRM_OK;
}

crm_status crm_get_snapshot(crm_ctx* ctx, double* records, int32_t* species, uint32_t* n_hist) {
    if (!ctx || !ctx->have_agents) return CRM_INVALID_ARGUMENT;
    uint32_t slot[2];
    const uint32_t nh = history_slots(ctx, slot);
    if (n_hist) *n_hist = nh;
    if (!records && !species) return CRM_OK;  // sizing call
    CRM_HIP(hipSetDevice(ctx->device));
    uint32_t nh2 = 0;
    size_t len = 0;
    CRM_TRY(gather_snapshot_device(ctx, nh2, len));
    if (records)
        CRM_HIP(hipMemcpyAsync(records, ctx->d_snap, len * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (species)
        CRM_HIP(hipMemcpyAsync(species, ctx->d_snap_sp, ctx->N * sizeof(int32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
    CRM_HIP(hipStreamSynchronize(ctx->stream));
    return CRM_OK;
}

crm_status crm_save_cells_json(crm_ctx* ctx, const char* path, uint64_t iteration, const crm_json_cells* spec) {
    if (!ctx || !path || !spec || !spec->cell_templates || spec->n_templates == 0) return CRM_INVALID_ARGUMENT;
    if (!ctx->have_agents) return CRM_INVALID_ARGUMENT;
    std::vector<Template> T;
    for (uint32_t i = 0; i < spec->n_templates; i++) {
        if (!spec->cell_templates[i]) {
            ctx->err = "cell template " + std::to_string(i) + " is NULL";
            return CRM_INVALID_ARGUMENT;
        }
        T.push_back(split_template(spec->cell_templates[i]));
    }
    CRM_HIP(hipSetDevice(ctx->device));
    std::vector<double> rec;
    std::vector<int32_t> sp;
    uint32_t nh = 0;
    CRM_TRY(gather_snapshot(ctx, rec, sp, nh));
    const uint64_t N = ctx->N;
    std::vector<double> rodv;
    std::vector<uint64_t> ids;
    std::vector<uint32_t> held;
    if (!is_rod(ctx)) {  // identifiers from the host copy; every agent holds the whole ring but daughters
        ids.resize(6 * N);
        held.assign(N, nh);
        if (ctx->pt_births) {
            CRM_HIP(hipMemcpyAsync(held.data(), ctx->d_snap_st, N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
            CRM_HIP(hipStreamSynchronize(ctx->stream));
        }
        for (uint64_t a = 0; a < N; a++)
            for (int q = 0; q < 3; q++) {
                ids[6 * a + q] = ctx->pt_ident[3 * a + q];
                ids[6 * a + 3 + q] = ctx->pt_parent[3 * a + q];
            }
    } else {
        rodv.resize(2 * N);
        ids.resize(6 * N);
        held.resize(N);
        CRM_HIP(hipMemcpyAsync(rodv.data(), ctx->d_snap_rod, rodv.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipMemcpyAsync(ids.data(), ctx->d_snap_id, ids.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipMemcpyAsync(held.data(), ctx->d_snap_st, N * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        CRM_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (spec->n_templates > 1)
        for (uint64_t a = 0; a < N; a++)
            if (sp[a] < 0 || (uint32_t)sp[a] >= spec->n_templates) {
                ctx->err = "agent " + std::to_string(a) + " has species " + std::to_string(sp[a]) +
                           " but only " + std::to_string(spec->n_templates) + " cell templates were given";
                return CRM_INVALID_ARGUMENT;
            }
    // <path>/<iteration, 20 digits>/batch_<instance, 20 digits>_<first free counter, 20 digits>.json
    // (storage/concepts.rs:702-740)
    namespace fs = std::filesystem;
    char name[96];
    std::snprintf(name, sizeof name, "%020llu", (unsigned long long)iteration);
    const fs::path dir = fs::path(path) / name;
    std::error_code ec;
    fs::create_directories(dir, ec);
    if (ec) {
        ctx->err = "cannot create " + dir.string() + ": " + ec.message();
        return CRM_STORAGE_ERROR;
    }
    fs::path file;
    for (uint64_t c = 0;; c++) {
        std::snprintf(name, sizeof name, "batch_%020llu_%020llu.json", (unsigned long long)spec->storage_instance,
                      (unsigned long long)c);
        file = dir / name;
        if (!fs::exists(file)) break;
    }
    // format in parallel over contiguous agent ranges
    const uint32_t nv = ctx->nv, D = ctx->D;
    const size_t rec_len = (size_t)ctx->K * (2u + 2u * nh);
    const uint32_t n_thr = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({16, N / 4096 + 1,
                                                          std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<std::string> part(n_thr);
    auto work = [&](uint32_t t) {
        const uint64_t a0 = N * t / n_thr, a1 = N * (t + 1) / n_thr;
        std::string& o = part[t];
        o.reserve((a1 - a0) * (400 + 60 * rec_len));
        for (uint64_t a = a0; a < a1; a++) {
            if (a) o += ",\n";
            put_cell(o, rec.data() + a * rec_len, T[spec->n_templates > 1 ? sp[a] : 0], nv, D, nh,
                     ids.data() + 6 * a, rodv.empty() ? nullptr : rodv.data() + 2 * a, held[a]);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < n_thr; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    FILE* f = std::fopen(file.string().c_str(), "wb");
    if (!f) {
        ctx->err = "cannot open " + file.string() + " for writing";
        return CRM_STORAGE_ERROR;
    }
    bool ok = std::fputs("{\n  \"data\": [\n", f) >= 0;
    for (const auto& p : part) ok = ok && std::fwrite(p.data(), 1, p.size(), f) == p.size();
    ok = ok && std::fputs("\n  ]\n}", f) >= 0;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        ctx->err = "write to " + file.string() + " failed";
        return CRM_STORAGE_ERROR;
    }
    return CRM_OK;
}

crm_status crm_get_tiling(crm_ctx* ctx, uint32_t inst, uint32_t* rows, uint32_t* cols, uint32_t* n_tiles,
                          uint32_t* lanes_per_agent) {
    if (!ctx || !ctx->have_agents || inst >= ctx->n_inst || is_rod(ctx)) return CRM_INVALID_ARGUMENT;
    if (rows) *rows = ctx->tile_rows[inst];
    if (cols) *cols = ctx->tile_cols[inst];
    if (n_tiles) *n_tiles = ctx->tile_count[inst];
    if (lanes_per_agent) *lanes_per_agent = 1u << ctx->lanes_log2;
    return CRM_OK;
}

crm_status crm_get_info(crm_ctx* ctx, crm_info* info) {
    if (!ctx || !info) return CRM_INVALID_ARGUMENT;
    std::memset(info, 0, sizeof(*info));
    info->n_agents = ctx->N;
    info->n_voxels_total = ctx->Vt;
    info->step = ctx->step;
    info->n_instances = ctx->n_inst;
    info->n_vertices = ctx->nv;
    info->dim = ctx->D;
    if (ctx->have_agents && ctx->d_ctrl) {
        Control c;
        CRM_TRY(read_control(ctx, c));
        if (c.error_flags) {
            info->error_step_valid = 1;
            info->error_step = c.error_key >> 32;
            info->error_agent = c.error_key & 0xffffffffu;
        }
    }
    return CRM_OK;
}

crm_status crm_get_timing(crm_ctx* ctx, crm_timing* timing) {
    if (!ctx || !timing) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_TRY(t_flush(ctx));
    for (int i = 0; i < CRM_KERNEL_COUNT; i++) {
        timing->total_ms[i] = ctx->t_ms[i];
        timing->launches[i] = ctx->t_n[i];
    }
    return CRM_OK;
}

crm_status crm_reset_timing(crm_ctx* ctx) {
    if (!ctx) return CRM_INVALID_ARGUMENT;
    CRM_HIP(hipSetDevice(ctx->device));
    CRM_TRY(t_flush(ctx));  // pairs still in flight belong to the timings cleared here
    for (int i = 0; i < CRM_KERNEL_COUNT; i++) {
        ctx->t_ms[i] = 0.0;
        ctx->t_n[i] = 0;
    }
    return CRM_OK;
}

size_t crm_last_error(crm_ctx* ctx, char* buf, size_t len) {
    const std::string& e = ctx ? ctx->err : g_create_error;
    if (buf && len) {
        size_t n = std::min(len - 1, e.size());
        std::memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return e.size();
}

void crm_destroy(crm_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    free_device(ctx);
    if (ctx->d_ip) (void)hipFree(ctx->d_ip);
    if (ctx->d_ctrl) (void)hipFree(ctx->d_ctrl);
    for (int i = 0; i < crm_ctx::kTimingPairs; i++) {
        if (ctx->ev_a[i]) (void)hipEventDestroy(ctx->ev_a[i]);
        if (ctx->ev_b[i]) (void)hipEventDestroy(ctx->ev_b[i]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"
