// kernels_remove.hpp -- removal of agents (CycleEvent::Remove): the flagged
// agents leave their voxel's list and the others keep their order
// (update_cycle.rs:52-55, Vec::retain); identifiers and the voxels' id counters
// are not touched.  Agent indices are compacted in order (agent a keeps the
// rank of a among the survivors).
//
//   k_remove_mark     per slot: survivor flag; per agent index: survivor flag;
//                     removed agents per instance
//   (exclusive scans of both flags)
//   k_remove_compact  survivors of the old order packed into the other buffer
//                     (same order, new agent index)
//   k_remove_key      voxel of every survivor and its arrival rank, then the
//                     counting sort of a set-up step rebuilds the voxel lists,
//                     the force records and the keys binned ahead
#pragma once

#include "kernels_rod.hpp"

namespace crm {

struct RemoveArgs {
    const uint8_t* flag;   // [n_old] by agent index: nonzero = remove
    uint32_t* keep_slot;   // [n_old + 1] 1 for a surviving slot (the last item 0)
    uint32_t* keep_id;     // [n_old + 1] 1 for a surviving agent index (the last item 0)
    uint32_t* slot_before; // [n_old + 1] exclusive scan of keep_slot
    uint32_t* id_before;   // [n_old + 1] exclusive scan of keep_id
    uint32_t* inst_rm;     // [n_inst] removed agents per instance
    uint64_t n_old;
    uint32_t n_inst;
};

// x[0 .. n) = v (clears counters without a memset node).
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* x, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = v;
}

// Thread q in [0, n_old]: q is both a slot (survivor flag by slot, and the
// instance count through the slot's voxel) and an agent index.
__global__ __launch_bounds__(kBlock) void k_remove_mark(Arrays A, StateBuf S, RemoveArgs X) {
    const uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q > X.n_old) return;
    if (q == X.n_old) {
        X.keep_slot[q] = 0u;
        X.keep_id[q] = 0u;
        return;
    }
    X.keep_id[q] = X.flag[q] ? 0u : 1u;
    const bool rm = X.flag[S.id[q]] != 0;
    X.keep_slot[q] = rm ? 0u : 1u;
    if (!rm) return;
    // the instance whose voxel range holds the agent's voxel: the last instance
    // whose first voxel is at or below it (voxel offsets ascend with the instance)
    const uint32_t g = S.vox[q];
    uint32_t lo = 0, hi = X.n_inst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A.ip[mid].voxel_offset <= g) lo = mid;
        else hi = mid;
    }
    const InstParams& P = A.ip[lo];
    if (g >= P.voxel_offset && g < P.voxel_offset + P.n_voxels) atomicAdd(&X.inst_rm[lo], 1u);
}

// Survivor in slot s of src (stride n_old) -> slot slot_before[s] of dst (stride
// n_new); K values per vector, `cap` history slots, rod extras when `rod`.
__global__ __launch_bounds__(kBlock) void k_remove_compact(StateBuf src, StateBuf dst, RemoveArgs X, uint64_t n_new,
                                                           uint32_t K, uint32_t cap, uint32_t rod) {
    const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t n_old = X.n_old;
    if (s >= n_old || !X.keep_slot[s]) return;
    const uint64_t m = X.slot_before[s];
    for (uint32_t c = 0; c < K; c++) {
        dst.pos[c * n_new + m] = src.pos[c * n_old + s];
        dst.vel[c * n_new + m] = src.vel[c * n_old + s];
        for (uint32_t h = 0; h < cap; h++) {
            dst.hdx[((uint64_t)h * K + c) * n_new + m] = src.hdx[((uint64_t)h * K + c) * n_old + s];
            dst.hdv[((uint64_t)h * K + c) * n_new + m] = src.hdv[((uint64_t)h * K + c) * n_old + s];
        }
    }
    dst.species[m] = src.species[s];
    dst.id[m] = X.id_before[src.id[s]];
    dst.vox[m] = src.vox[s];
    if (rod) copy_rod_extras(src, s, dst, m, n_old, n_new);
    else if (src.birth) dst.birth[m] = src.birth[s];  // point agents inserted by a division
}

// Next order of the compacted agents: each stays in its voxel; k_fixup orders
// every voxel by slot, so the arrival ranks need no order.
__global__ __launch_bounds__(kBlock) void k_remove_key(Arrays A, StateBuf S) {
    const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= A.N) return;
    const uint32_t key = S.vox[s];
    A.key[s] = key;
    A.rank[s] = atomicAdd(&A.count[key], 1u);
}

}  // namespace crm
