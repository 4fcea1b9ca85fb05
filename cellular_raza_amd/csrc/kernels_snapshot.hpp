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
    double* rod;          // rods (may be null): [N][2] spring length, growth rate
    uint64_t* ident;      // rods (may be null): [N][6] identifier (kind, a, b), parent (kind, a, b)
    uint32_t* stored;     // rods (may be null): [N] history entries the rod holds (a daughter holds fewer)
    uint64_t step;        // completed steps
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
    if (a.rod) {
        a.rod[2ull * id] = S.spring[s];
        a.rod[2ull * id + 1] = S.growth[s];
    }
    if (a.stored) {
        const uint64_t pushes = a.step > S.birth[s] ? a.step - S.birth[s] : 0;
        a.stored[id] = (uint32_t)min<uint64_t>(pushes, a.n_hist);
    }
    if (a.ident) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            a.ident[6ull * id + j] = S.ident[(uint64_t)j * N + s];
            a.ident[6ull * id + 3 + j] = S.parent[(uint64_t)j * N + s];
        }
    }
}

}  // namespace crm
