/*
 * crmech.h -- C ABI of the MI355X (gfx950) mechanics backend for cellular_raza.
 *
 * The library advances agent-based mechanics simulations on one GPU:
 *   voxel binning on a CartesianCuboid grid (integer counting sort by voxel),
 *   neighbour search over the 3^D voxel stencil,
 *   the pairwise Interaction::calculate_force_between loop,
 *   the Mechanics update (NewtonDamped / Langevin / Brownian / rods) with the
 *   chili Euler / Adams-Bashforth steppers, reflecting boundary, re-binning.
 *
 * Reference interfaces this ABI replaces (paths relative to the cellular_raza
 * workspace):
 *   - the per-step hook functions that run_simulation! calls,
 *     cellular_raza-core-proc-macro/src/run_sim.rs:152-157,251-256
 *     (update_mechanics_interaction_step_{1,2,3}), 259-270 (local mechanics
 *     update), 293-297 (apply_boundary, sort_cells_in_voxels_step_{1,2});
 *   - the native-backend FFI precedent cellular_raza-core/src/backend/cara/mod.rs:5-62
 *     (opaque runner, extern "C" create/drop/update) and its build.rs:7-19.
 *
 * Conventions
 *   - All host arrays are caller-owned and copied; the context owns device memory.
 *   - "Agent index" is the position of an agent in the arrays passed to
 *     crm_set_agents (instances concatenated in order).  Every array the library
 *     returns is in agent-index order.
 *   - Vectors are row-major: agent a, vertex v, component k at
 *     ((a * n_vertices) + v) * dim + k.  Point agents have n_vertices == 1.
 *   - Calls are synchronous with respect to the context's HIP stream.  One
 *     context must not be used from two host threads at once.
 *   - Status codes mirror cellular_raza-concepts/src/errors.rs:4-25,63-101 and
 *     cellular_raza-core/src/backend/chili/errors.rs:73-121.
 */
#ifndef CRMECH_H
#define CRMECH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRM_MAX_DIM 3
#define CRM_MAX_SPECIES 4
#define CRM_MAX_ROD_VERTICES 16

typedef enum crm_status {
    CRM_OK = 0,
    CRM_CALC_ERROR = 1,       /* CalcError      (errors.rs: e.g. Mie at r == 0, interaction.rs:355-359) */
    CRM_BOUNDARY_ERROR = 2,   /* BoundaryError  (cartesian_cuboid_n.rs:149-162,615-625) */
    CRM_INDEX_ERROR = 3,      /* IndexError     (update_mechanics.rs:433-451; initial decompose) */
    CRM_RNG_ERROR = 4,        /* RngError       (mechanics.rs:211-226) */
    CRM_SETUP_ERROR = 5,      /* DecomposeError / TimeError during setup */
    CRM_DEVICE_ERROR = 6,     /* HIP runtime failure */
    CRM_INVALID_ARGUMENT = 7, /* caller error: bad sizes, null pointers, unsupported combination */
    CRM_STORAGE_ERROR = 8,    /* StorageError (storage/concepts.rs:16-42): a snapshot file could not be written */
    CRM_DIVISION_ERROR = 9    /* DivisionError (concepts errors.rs): RodMechanics::divide failed */
} crm_status;

/* Interaction kinds (cellular_raza-building-blocks/src/cell_building_blocks/interaction.rs). */
typedef enum crm_potential_kind {
    CRM_POTENTIAL_MORSE = 0,              /* MorsePotential, interaction.rs:151-299 */
    CRM_POTENTIAL_MIE = 1,                /* MiePotential, interaction.rs:301-434 */
    CRM_POTENTIAL_BOUND_LENNARD_JONES = 2,/* BoundLennardJones, interaction.rs:84-147 */
    CRM_POTENTIAL_MORSE_SPECIES = 3,      /* Morse with strength * species_matrix[own][ext];
                                             composite, no single reference type */
    CRM_POTENTIAL_CELL_SORTING = 4,       /* CellSpecificInteraction of
                                             cellular_raza-benchmarks/src/cell_sorting.rs:31-94 (sign = +1)
                                             and cellular_raza-examples/cell_sorting/src/main.rs:32-103 (sign = -1) */
    CRM_POTENTIAL_NONE = 5,               /* NoInteraction, interaction.rs:12-35 (Mechanics-only runs) */
    CRM_POTENTIAL_CELL_SORTING_LANGEVIN = 6 /* CellSpecificInteraction of
                                             cellular_raza-examples/cell_sorting/src/main-langevin.rs:44-92:
                                             repulsion (strength) * (R_o + R_e - r) in contact, Gaussian-tail
                                             attraction species_matrix[own][ext] up to cutoff, (-force, force) */
} crm_potential_kind;

/* Mechanics kinds (cellular_raza-building-blocks/src/cell_building_blocks/mechanics.rs,
 * bacterial_rods.rs). */
typedef enum crm_mechanics_kind {
    CRM_MECHANICS_NEWTON_DAMPED = 0,  /* NewtonDamped{2,3}D, mechanics.rs:14-204 */
    CRM_MECHANICS_LANGEVIN = 1,       /* Langevin{2,3}D, mechanics.rs:389-571 */
    CRM_MECHANICS_BROWNIAN = 2,       /* Brownian{2,3}D, mechanics.rs:228-387 */
    CRM_MECHANICS_ROD = 3             /* RodMechanics + RodInteraction, bacterial_rods.rs:82-337 */
} crm_mechanics_kind;

typedef enum crm_precision {
    CRM_PRECISION_FP64 = 0,   /* everything in double */
    CRM_PRECISION_MIXED = 1   /* float pair forces on voxel-relative coordinates and
                                 float per-agent sums; double state and integrator */
} crm_precision;

typedef enum crm_weighting {
    /* chili with one subdomain: same-voxel pairs get 1/2 of each side's evaluation,
     * adjacent-voxel pairs only 1/2 of the own evaluation
     * (update_mechanics.rs:178-220, 63-109). */
    CRM_WEIGHT_REFERENCE = 0,
    /* every pair in the stencil weighted like a same-voxel pair (full force),
     * which is what chili gives pairs that straddle two subdomains (update_mechanics.rs:285-382). */
    CRM_WEIGHT_PHYSICAL = 1
} crm_weighting;

typedef enum crm_history_order {
    CRM_HISTORY_OLDEST_FIRST = 0,  /* ring-buffer iteration oldest first (aux_storage.rs:859-877) */
    CRM_HISTORY_NEWEST_FIRST = 1
} crm_history_order;

typedef enum crm_noise_mode {
    CRM_NOISE_NONE = 0,     /* deterministic mechanics only (noise terms are zero) */
    CRM_NOISE_BUFFER = 1,   /* caller passes standard normals per step, by agent index */
    CRM_NOISE_PHILOX = 2    /* on-device Philox4x32-10 + Box-Muller keyed by (seed, agent index, step) */
} crm_noise_mode;

enum {
    CRM_FLAG_GRAPH = 1u,    /* replay the per-step launch sequence as a captured hipGraph */
    CRM_FLAG_TIMING = 2u,   /* record HIP events around every kernel (disables graphs) */
    CRM_FLAG_NO_LDS = 4u    /* force kernel reads candidates from global memory (testing) */
};

/* CartesianCuboid, cartesian_cuboid_n.rs:65-72,183-256.
 * n_voxels[0] == 0 selects from_boundaries_and_interaction_range, otherwise
 * from_boundaries_and_n_voxels. */
typedef struct crm_domain {
    double min[3];
    double max[3];
    double interaction_range;
    uint32_t n_voxels[3];
} crm_domain;

typedef struct crm_potential {
    uint32_t kind;                    /* crm_potential_kind */
    uint32_t n_species;               /* MORSE_SPECIES / CELL_SORTING / CELL_SORTING_LANGEVIN: species are
                                         0..n_species-1 */
    double radius;                    /* Morse / Mie radius; CellSpecificInteraction cell_radius */
    double potential_stiffness;       /* Morse lambda */
    double cutoff;                    /* Morse / Mie / BoundLennardJones / CellSortingLangevin cutoff */
    double strength;                  /* Morse V0, Mie epsilon ("strength"), CellSorting potential_strength,
                                         CellSortingLangevin repulsion */
    double bound;                     /* Mie / BoundLennardJones bound */
    double en, em;                    /* Mie exponents */
    double epsilon, sigma;            /* BoundLennardJones */
    double relative_interaction_range;/* CellSorting */
    double sign;                      /* CellSorting: +1 benchmark variant, -1 example variant */
    double species_matrix[CRM_MAX_SPECIES * CRM_MAX_SPECIES]; /* MORSE_SPECIES: strength factor;
                                                                 CELL_SORTING_LANGEVIN: attraction
                                                                 (rr, rb / br, bb); row = own species */
} crm_potential;

typedef struct crm_mechanics {
    uint32_t kind;             /* crm_mechanics_kind */
    uint32_t n_vertices;       /* rods: vertices per rod (2..16); points: 1 */
    double mass;               /* NewtonDamped / Langevin */
    double damping;            /* NewtonDamped damping_constant, Langevin damping, Rod damping */
    double kb_temperature;     /* Langevin / Brownian */
    double diffusion_constant; /* Brownian / Rod */
    double spring_tension;     /* Rod */
    double rigidity;           /* Rod */
    double spring_length;      /* Rod */
} crm_mechanics;

/* One independent simulation.  Several instances in one context advance together
 * (ensemble batching); they share dim, kinds, dt and solver, and may differ in
 * domain and parameters. */
typedef struct crm_instance {
    crm_domain domain;
    crm_potential potential;
    crm_mechanics mechanics;
} crm_instance;

typedef struct crm_config {
    uint32_t dim;            /* 2 or 3 */
    uint32_t solver_order;   /* run_simulation! mechanics_solver_order: 0 Euler (solvers.rs:23-55),
                                1 mechanics_adams_bashforth_2 (247-300), 2 mechanics_adams_bashforth_3
                                (156-231, the default, run_sim.rs:92) */
    uint32_t precision;      /* crm_precision */
    uint32_t weighting;      /* crm_weighting */
    uint32_t history_order;  /* crm_history_order */
    uint32_t noise_mode;     /* crm_noise_mode */
    uint64_t seed;
    double dt;
    int32_t device;
    uint32_t flags;          /* CRM_FLAG_* */
    /* CRM_WEIGHT_REFERENCE only: chili's subdomain count (Settings::n_threads,
     * chili/settings.rs; create_subdomains, cartesian_cuboid_n.rs:19-59,493-584).
     * 0 or 1: one subdomain.  k > 1: every instance's voxels are split into k
     * lexicographic chunks by get_decomp_res, and a pair of agents in adjacent
     * voxels owned by different chunks gets the full force (both halves, as
     * update_mechanics_interaction_step_2/3 deliver them, update_mechanics.rs:
     * 285-382) instead of 1/2 of the own evaluation.  Point agents only. */
    uint32_t n_subdomains;
    uint32_t reserved;
} crm_config;

typedef struct crm_info {
    uint64_t n_agents;
    uint64_t n_voxels_total;     /* sum over instances */
    uint64_t step;               /* steps executed so far */
    uint32_t n_instances;
    uint32_t n_vertices;
    uint32_t dim;
    uint32_t error_step_valid;
    uint64_t error_step;         /* first step that raised an error */
    uint64_t error_agent;        /* agent index that raised it */
} crm_info;

/* Kernel classes reported by crm_get_timing, in this order. */
enum {
    CRM_KERNEL_FORCE = 0,
    CRM_KERNEL_INTEGRATE = 1,
    CRM_KERNEL_SCAN = 2,
    CRM_KERNEL_SCATTER = 3,
    CRM_KERNEL_FIXUP = 4,
    CRM_KERNEL_GATHER = 5,
    CRM_KERNEL_COUNT = 6
};

typedef struct crm_timing {
    double total_ms[CRM_KERNEL_COUNT];  /* summed HIP-event time per kernel class */
    uint64_t launches[CRM_KERNEL_COUNT];
} crm_timing;

typedef struct crm_ctx crm_ctx;

/* Grid of a CartesianCuboid (cartesian_cuboid_n.rs:183-256); no device work. */
crm_status crm_grid(const crm_domain* domain, uint32_t dim, uint64_t n_voxels_out[3], double dx_out[3]);

/* Creates a context on cfg->device.  Replaces construct_simulation_runner
 * (chili/datastructures.rs:51-217) for the mechanics aspects.  On failure *out
 * is NULL and crm_last_error(NULL, ...) holds the reason. */
crm_status crm_create(const crm_config* cfg, const crm_instance* instances, uint32_t n_instances,
                      crm_ctx** out);

/* Uploads agents and bins them (Domain::decompose + insert_initial_cells,
 * cellular_raza-concepts-derive/src/domain.rs:292-421, datastructures.rs:278-299).
 * counts[i] agents belong to instance i.  pos, vel: sum(counts) * n_vertices * dim doubles.
 * vel may be NULL (zero).  species may be NULL (all 0).  Resets the step counter
 * and the Adams-Bashforth history. */
crm_status crm_set_agents(crm_ctx* ctx, const uint64_t* counts, const double* pos, const double* vel,
                          const int32_t* species);

/* Advances n_steps steps: forces -> integrator -> boundary -> re-binning
 * (run_sim.rs:380-403).  noise: CRM_NOISE_BUFFER only, n_steps * n_agents *
 * n_vertices * dim standard normals, step-major, by agent index.  A step that
 * fails (CalcError, BoundaryError, DivisionError) ends, and no later step runs:
 * the call returns the error, crm_get_info reports the failing step and agent,
 * the step counter and the state (positions, history, voxel lists) are those at
 * the failing step's end, and the context accepts no further step. */
crm_status crm_step(crm_ctx* ctx, uint32_t n_steps, const double* noise);

/* Interaction force accumulated for the current state, by agent index
 * (AuxStorageMechanics::current_force before integration, aux_storage.rs:41-130).
 * Does not advance the simulation.  Point agents in mixed precision hold their
 * forces in float (the values returned, and the ones the integrator uses). */
crm_status crm_compute_forces(crm_ctx* ctx, double* force);

crm_status crm_get_state(crm_ctx* ctx, double* pos, double* vel);

/* Adams-Bashforth ring contents, oldest slot first: dx_hist, dv_hist each
 * 2 * n_agents * n_vertices * dim doubles (slot-major); n_stored in {0,1,2}.  A
 * point agent inserted by crm_insert_agents holds only the pushes since its
 * insertion (crm_get_identifiers' history_held): they come first, zeros after. */
crm_status crm_get_history(crm_ctx* ctx, double* dx_hist, double* dv_hist, uint32_t* n_stored);

/* Voxel of every agent (instance voxel offset + plain index, plain index as in
 * datastructures.rs:113-120) and the CSR offsets of the sorted agent list
 * (n_voxels_total + 1 entries); either pointer may be NULL. */
crm_status crm_get_voxels(crm_ctx* ctx, uint32_t* voxel_of_agent, uint32_t* voxel_start);

/* The voxels' cell lists: order[n_agents] holds agent indices in sorted order, so
 * order[voxel_start[v] .. voxel_start[v + 1]) lists voxel v's agents in the order
 * of the reference's Voxel::cells (datastructures.rs:349-396): after each step the
 * agents that stayed keep their order and the arrivals follow by old voxel, then
 * old position (sort_cells_in_voxels, update_mechanics.rs:418-513); daughters of
 * the rods cycle follow their voxel's list (update_cycle.rs:57-66). */
crm_status crm_get_voxel_order(crm_ctx* ctx, uint32_t* order);

crm_status crm_get_info(crm_ctx* ctx, crm_info* info);

/* Tiling autotuner (no counterpart in the reference, whose CPU backend has no
 * work tiles): times the force kernel on the current agents under each of the
 * library's tiling policies and keeps the fastest.  A policy sets the voxel tile
 * shape (through the own agents per lane of a 256-lane workgroup, up to 1.9 with
 * a lane walking several agents in turn, and the density margin of the staged
 * region) and the lanes per agent (1, 2, 4 or 8: 64, 32, 16 or 8 agents per
 * wavefront; the lanes of an agent split its candidate stream and add their
 * partial forces with wavefront shuffles).  The
 * choice is cached per workload key (dimension, precision, potential, weighting,
 * instance count, agent count, grid) for the process and, when cache_path is not
 * NULL, in that JSON file, which is read first and rewritten.  FP64 forces do
 * not depend on the tile shape, and with more than one lane per agent differ only
 * by the association of the partial sums (last bits); mixed-precision forces
 * quantise positions with the tile region's fixed-point step and so move with the
 * tiling at that step's level (reproducible for one tiling).  chosen (may be
 * NULL) receives the policy index. */
crm_status crm_autotune(crm_ctx* ctx, const char* cache_path, uint32_t* chosen);

/* Force tiling of instance inst: voxel rows and columns per tile, tile count,
 * lanes per agent (any output may be NULL). */
crm_status crm_get_tiling(crm_ctx* ctx, uint32_t inst, uint32_t* rows, uint32_t* cols, uint32_t* n_tiles,
                          uint32_t* lanes_per_agent);

/* Force of the domain on every agent (the DomainForce aspect: SubDomainForce::
 * calculate_custom_force, concepts domain.rs:250-285, applied per cell after the
 * subdomain's pair forces by calculate_custom_domain_force,
 * update_mechanics.rs:244-274, run_sim.rs:277-279).  Per point or rod vertex:
 *   SPRING    f_k = (-strength[k]) * (x_k - center[k])   (the form of the
 *             reference's test domain, cellular_raza/tests/domain_force_spring.rs:56-67)
 *   ROD_GEL   CartesianSubDomainRods (bacterial_rods.rs:437-479), 3-D rods only:
 *             (0, 0, -gel_pressure), plus -(v / |v|) * gel_pressure *
 *             surface_friction when |v| != 0 and the vertex is within
 *             surface_friction_distance of a domain face on any axis.
 * Vectors are in the caller's axis order (x, y[, z]). */
typedef enum crm_domain_force_kind {
    CRM_DOMAIN_FORCE_NONE = 0,
    CRM_DOMAIN_FORCE_SPRING = 1,
    CRM_DOMAIN_FORCE_ROD_GEL = 2
} crm_domain_force_kind;

typedef struct crm_domain_force {
    uint32_t kind;
    uint32_t pad;
    double strength[3];
    double center[3];
    double gel_pressure;
    double surface_friction;
    double surface_friction_distance;
} crm_domain_force;

/* Sets the domain force of instance `inst` (default NONE). */
crm_status crm_set_domain_force(crm_ctx* ctx, uint32_t inst, const crm_domain_force* force);

/* Cycle of the rods example (aspects [Mechanics, Interaction, Cycle,
 * NeighborSensing], bacterial_rods/src/main.rs:48-111) with chili's division
 * bookkeeping (update_cycle.rs:10-69).  Every step, after the mechanics update:
 *   react_to_neighbors: n = vertex pairs of the stencil closer than
 *     sensing_radius (the NeighborSensing count of this step's force pass);
 *     growth_rate = ((growth_base_rate * (neighbor_saturation - n)) /
 *     neighbor_saturation).max(0) if n > 0, else growth_base_rate;
 *   update_cycle: spring_length += growth_rate * dt; Division when
 *     spring_length > division_threshold;
 *   update_cell_cycle_4: RodMechanics::divide(division_radius) (bacterial_rods.rs:
 *     586-662) splits the mother in place and adds a daughter (a clone with the
 *     other half, empty history); identifiers CellIdentifier::Division(voxel,
 *     counter) from the voxel's counter, mothers first, the parent the mother's
 *     old identifier;
 * then the boundary and the re-sort.  Daughters get agent indices n_agents,
 * n_agents + 1, ... in voxel order.  The context holds at most max_agents rods
 * (CRM_INVALID_ARGUMENT past it); rods only.  Enabling it switches neighbour
 * sensing on with sensing_radius and sets every growth_rate to the base rate. */
typedef struct crm_rod_cycle {
    double growth_base_rate;
    double division_threshold;
    double division_radius;
    double sensing_radius;
    double neighbor_saturation;
    uint64_t max_agents;
} crm_rod_cycle;
crm_status crm_set_rod_cycle(crm_ctx* ctx, const crm_rod_cycle* cycle);

/* Per-rod spring lengths (RodMechanics::spring_length, bacterial_rods.rs:104), by
 * agent index; crm_set_agents initialises them from the instance's mechanics. */
crm_status crm_set_rod_spring_lengths(crm_ctx* ctx, const double* spring_length);

/* Per-rod cycle state by agent index (any pointer may be NULL): spring_length[N],
 * growth_rate[N], identifier[3 N] and parent[3 N] as (kind, a, b) with kind 0
 * Initial(a), 1 Division(a = voxel plain index, b = counter), 3 no parent, and
 * history_held[N], the AuxStorageMechanics entries the rod holds (a daughter
 * starts empty): its newest entries of crm_get_history's. */
crm_status crm_get_cycle_state(crm_ctx* ctx, double* spring_length, double* growth_rate, uint64_t* identifier,
                               uint64_t* parent, uint32_t* history_held);

/* CellIdentifier and parent of every agent by agent index (cellular_raza-concepts
 * cell.rs:18-25, CellBox::identifier / parent, cell.rs:128-137), points and rods,
 * each as (kind, a, b): kind 0 Initial(a) (a = the agent's index among its
 * instance's initial cells, datastructures.rs:286-297), 1 Division(a = voxel plain
 * index, b = counter), 3 no parent; history_held[N] = the AuxStorageMechanics
 * entries each agent holds.  identifier and parent take 3 * n_agents values; any
 * pointer may be NULL.  Identifiers survive removals (crm_remove_agents
 * renumbers only the agent indices). */
crm_status crm_get_identifiers(crm_ctx* ctx, uint64_t* identifier, uint64_t* parent, uint32_t* history_held);

/* Removes the agents with remove[a] != 0 (remove[n_agents], by agent index):
 * CycleEvent::Remove as update_cell_cycle_4 applies it (chili/update_cycle.rs:
 * 41-55: the flagged cells leave their voxel's list, Vec::retain keeps the order of
 * the others, and neither identifiers nor the voxels' id counters change).  Call
 * it between steps for the agents whose cycle returned Remove in the step just
 * taken (or whose update_conditional_phased_death returned true, update_cycle.rs:
 * 113-119): the reference removes them after that step's mechanics update and
 * before its sort, which leaves the same voxel lists.  The survivors keep their
 * state, history and cycle state; agent indices are compacted in order (agent a
 * becomes the number of survivors before a), which also renumbers the Philox noise
 * streams and the rows of every by-agent array.  Removing every agent is
 * CRM_INVALID_ARGUMENT.  Points and rods, with or without the rods cycle.  The
 * context keeps the call's scratch memory (device allocation on the first call
 * and when the agent count exceeds every earlier call's). */
crm_status crm_remove_agents(crm_ctx* ctx, const uint8_t* remove);

/* CycleEvent::Division of a caller's Cycle implementation for point agents
 * (update_cycle.rs:28-66; the reference calls C::divide and appends the new cells
 * in Voxel::update_cell_cycle_4).  Call it between steps for the agents whose
 * update_cycle returned Division in the step just taken, with the states C::divide
 * produced: mother_pos / mother_vel (n * dim doubles each, NULL = unchanged) replace
 * mother[i]'s state, which keeps its history; daughter i (daughter_pos, and
 * daughter_vel or, if NULL, the mother's new velocity; daughter_species or, if
 * NULL, the mother's species) becomes agent index N + i with an empty AuxStorage
 * (no history: its first step is an Euler step, as in the reference).
 * Identifiers: voxel by voxel in list order, each mother takes
 * Division(voxel, ++id_counter) and its old identifier as parent; then the
 * voxel's daughters, in the mothers' order, take Division(voxel, ++id_counter) and
 * their mother's old identifier as parent, and join the end of the voxel's list.
 * Then apply_boundary reflects mothers and daughters and the re-binning of a step
 * moves those now outside their list's voxel (arrivals after stayers, by old
 * voxel and old position).  The reference divides after a step's mechanics update
 * and before its boundary and re-binning; here the step's re-binning has already
 * moved the mothers, so a mother that changed voxel in that step divides in its
 * new voxel, and a daughter joins its voxel after that step's arrivals.  A mother
 * listed twice or not an agent, a species out of range: CRM_INVALID_ARGUMENT; a
 * state that apply_boundary cannot bring into the domain: CRM_BOUNDARY_ERROR;
 * nothing changes in either case.  Point agents only (rods divide through
 * crm_set_rod_cycle).  Costs one host round trip of the state (device
 * reallocation when the context's agent capacity is exceeded). */
crm_status crm_insert_agents(crm_ctx* ctx, uint64_t n, const uint64_t* mother, const double* mother_pos,
                             const double* mother_vel, const double* daughter_pos, const double* daughter_vel,
                             const int32_t* daughter_species);

/* No counterpart in the reference: crm_remove_agents for callers that cannot
 * allocate device memory (code under HIP graph capture elsewhere in the process,
 * tight serving loops, harnesses that forbid hipMalloc).  Same semantics and
 * results as crm_remove_agents, but its scratch (removal flags, survivor flags
 * by slot and by agent index with their exclusive scans, removed agents per
 * instance, the scans' temporary storage) lives in one static device array of
 * about 69 MiB per GPU, held until the process ends, instead of the context's
 * grow-only allocation: it calls no hipMalloc / hipFree.  Limits: at most
 * 4,194,304 agents in at most 65,536 instances, else CRM_INVALID_ARGUMENT with
 * nothing changed.  Every call on one GPU uses that GPU's array: calls that
 * overlap in time on one GPU (two host threads, or two contexts on that device
 * driven concurrently) race and must be serialised by the caller.  Graphs
 * captured before the call are released by the next crm_step, not here. */
crm_status crm_remove_agents_noalloc(crm_ctx* ctx, const uint8_t* remove);

/* Neighbour sensing fused into the force pass.  Replaces the NeighborSensing
 * hooks of the chili pair loop (concepts interaction.rs:64-79 trait,
 * update_mechanics.rs:125-139 neighbor_sensing_single, called per pair at
 * update_mechanics.rs:105 and 199-202) for the accumulator of the rods example
 * (bacterial_rods/src/main.rs:48-78): during every force evaluation each agent
 * counts, over the other agents of its voxel and of the adjacent voxels, the
 * (own point, other point) pairs -- points, or rod vertices -- with
 * norm(own - other) < radius.  enable = 0 switches it off; a NaN radius is
 * rejected; radius <= 0 counts nothing.  The setting survives crm_set_agents. */
crm_status crm_set_neighbor_sensing(crm_ctx* ctx, uint32_t enable, double radius);

/* Counts of the last force evaluation (the last step's, or crm_compute_forces'),
 * by agent index: counts[n_agents].  Their positions are the state before that
 * step's integration, as the reference's accumulator sees it when
 * react_to_neighbors runs (update_mechanics.rs:561-576).  CRM_INVALID_ARGUMENT
 * when no force evaluation with sensing on has run since it was enabled or the
 * agents were set. */
crm_status crm_get_neighbor_counts(crm_ctx* ctx, uint32_t* counts);

/* Snapshot of the current state in agent-index order, gathered on the device in
 * one kernel and one copy.  Replaces the per-save-point walk of every voxel's
 * cells in SubDomainBox::save_cells (datastructures.rs:371-396).  records[N *
 * K * (2 + 2 * n_hist)] (may be NULL), K = n_vertices * dim, per agent:
 *   pos[K], vel[K], dx history[n_hist][K], dv history[n_hist][K]
 * with the history oldest first (AuxStorageMechanics positions / velocities
 * rings, aux_storage.rs:41-46,107-115); species[N] (may be NULL); n_hist
 * receives the number of stored history entries (0..2).  With records and
 * species both NULL only n_hist is returned (to size records). */
crm_status crm_get_snapshot(crm_ctx* ctx, double* records, int32_t* species, uint32_t* n_hist);

/* Cells of one save point written in the reference's JSON storage layout
 * (JsonStorageInterface, storage/serde_json.rs:10-74, StorageWrapper::
 * store_batch_elements and the file naming of storage/concepts.rs:702-790):
 *   <path>/<iteration:020>/batch_<storage_instance:020>_<n:020>.json
 * with n the first free counter, holding {"data": [CombinedSaveFormat ...]} with
 * one entry per agent: identifier (crm_get_identifiers), element [CellBox
 * {identifier, parent, cell}, AuxStorage {mechanics: {positions,
 * velocities, current_force, zero_force}}].  `cell` is the caller's JSON text
 * of the agent struct with the placeholders {pos} and {vel}, one template per
 * species (n_templates > 1) or one for all; points write vectors as [x, y, z],
 * rods as nalgebra's dynamic matrix [column-major data, rows, columns].  path is
 * the cells' json folder (<storage location>/cells/json). */
typedef struct crm_json_cells {
    const char* const* cell_templates;
    uint32_t n_templates;
    uint32_t storage_instance;
} crm_json_cells;
crm_status crm_save_cells_json(crm_ctx* ctx, const char* path, uint64_t iteration, const crm_json_cells* spec);

/* Kernel timing of a context created with CRM_FLAG_TIMING (no counterpart in the
 * reference).  The event pairs around the launches are recorded without waiting;
 * crm_get_timing waits for those still in flight and adds them in, and
 * crm_reset_timing does the same before clearing the sums.  A class may cover
 * several launches per step (the scans, or all three sort kernels when the voxels
 * fit one scan tile, count as CRM_KERNEL_SCAN). */
crm_status crm_get_timing(crm_ctx* ctx, crm_timing* timing);
crm_status crm_reset_timing(crm_ctx* ctx);

/* Copies the last error message (NUL-terminated) and returns its full length.
 * ctx == NULL reports why the last crm_create on this thread failed. */
size_t crm_last_error(crm_ctx* ctx, char* buf, size_t len);

void crm_destroy(crm_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* CRMECH_H */
