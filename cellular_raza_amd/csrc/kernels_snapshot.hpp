// Snapshot gather: the device state (SoA by sorted slot) -> one record per agent
// in agent-index order, so a save point costs one kernel and one device-to-host
// copy.  Record of agent a (doubles, K = n_vertices * dim):
//   [pos K][vel K][dx history n_hist x K][dv history n_hist x K]
// with the history oldest first -- the content of the reference's CellBox
// mechanics (pos, vel) and AuxStorageMechanics rings (positions, velocities:
// the increments pushed each step), aux_storage.rs:41-46,88-121.
#pragma once

#include "kernels_point.hpp"

namespace crm {

struct SnapshotArgs {
    double* out;          // [N][rec]
    int32_t* species;     // [N] agent-order species (may be null)
    uint64_t N;
    uint32_t K, n_hist;   // values per vector, stored history entries (0..2)
    uint32_t slot[2];     // ring slots of the history entries, oldest first
};

__global__ __launch_bounds__(kBlock) void k_snapshot(StateBuf S, SnapshotArgs a) {
    const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.N) return;
    const uint64_t N = a.N;
    const uint32_t K = a.K, rec = K * (2u + 2u * a.n_hist);
    const uint32_t id = S.id[s];
    double* o = a.out + (uint64_t)id * rec;
    for (uint32_t k = 0; k < K; k++) o[k] = S.pos[(uint64_t)k * N + s];
    for (uint32_t k = 0; k < K; k++) o[K + k] = S.vel[(uint64_t)k * N + s];
    const uint64_t ring = (uint64_t)N * K;  // elements per ring slot
    for (uint32_t i = 0; i < a.n_hist; i++) {
        double* dx = o + 2u * K + i * K;
        double* dv = o + 2u * K + (a.n_hist + i) * K;
        for (uint32_t k = 0; k < K; k++) {
            dx[k] = S.hdx[a.slot[i] * ring + (uint64_t)k * N + s];
            dv[k] = S.hdv[a.slot[i] * ring + (uint64_t)k * N + s];
        }
    }
    if (a.species) a.species[id] = S.species[s];
}

}  // namespace crm
