/* C-ABI of libbm_hip.so: the BladeMaster TLMPM cutting substep on MI355X (gfx950).
 *
 * The reference implements this path as Warp kernels launched from Python
 * (blademaster_warp/solver.py:318-940); its "FFI" is wp.launch(kernel, dim, args).  Each
 * entry point below replaces one such launch group and cites it.  All device buffers are
 * owned by the caller (the Python host keeps them as PyTorch-ROCm tensors) and passed as
 * plain pointers in bm_state; scalars travel in bm_params.  Every call is asynchronous on
 * the given stream (a hipStream_t passed as void*), never allocates, never synchronises,
 * and is safe to record into a HIP graph.  Stages that the reference guards with Warp
 * conditional graph nodes (capture_if on topology_dirty / pair_list_dirty) read the
 * corresponding device flag themselves and do nothing while it is zero.
 *
 * Return value: 0 on success, negative for an argument error (null pointer, bad size),
 * positive hipError_t for a launch failure.  Physics failures are counted on the device in
 * errors[3] = [storage overflow, inverted material, coincident proxies], the layout of
 * solver.py:307.
 *
 * Layout (float32 unless noted; n particles, N grid nodes, F field slots per node):
 *   particle vec3 arrays  [n][3], matrices [n][3][3] row-major, codes int64[n]
 *   local   [n][3] cell fractions of the reference position, base_node int32[n]
 *   field_idx uint8[n][8]  field slot of each stencil corner
 *   field_code int64[F][N] (-1 = empty slot; used slots form a prefix), node_nf int32[N]
 *   grid scalars [F][N], grid vectors [F][N][3]
 *   histories [2][n][3]  (row 0 board, row 1 blade friction history)
 *   W = (brick[0] + 1) (brick[1] + 1) (brick[2] + 1) window nodes per brick
 */
#ifndef BM_HIP_H
#define BM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bits of cut_bits (contact.py:15-18). */
#define BM_CUT_ON 1
#define BM_CUT_ADVANCE 2
#define BM_CUT_HOLD 4

typedef struct bm_params {
  int64_t n;              /* particles */
  int64_t node_count;     /* N = prod(dims) */
  int64_t fields;         /* F field slots per node */
  int64_t dims[3];        /* grid nodes per axis (cells + 1) */
  int64_t pair_capacity;  /* slots per pair-history table, power of two */
  int64_t list_buckets;   /* cell-list buckets, power of two */
  int64_t list_lanes;     /* listed partners per particle (16) */
  int64_t tool_model;     /* -1 external, 0 guided translation, 1 free rigid / servo */
  int64_t free_blade;     /* tool.step: no drive */
  int64_t lock_rotation;  /* tool.step: guided, translation only */
  int64_t predict;        /* finish_external_step: mobility predictor */
  int64_t plastic;        /* yield_stress > 0 */
  int64_t brick[3];       /* cells per brick along x, y, z (particle bins of the P2G/G2P kernels) */
  int64_t bricks[3];      /* bricks along x, y, z */
  int64_t rigid_blocks;   /* rows of rigid_partial: at least one per contact-kernel workgroup
                             (ceil(n / 256)); a substep needs 16 more, where its pair kernel
                             reduces the rows to 16 */
  int64_t blade_hold;     /* external tool held over the interval: grid constraint active */
  int64_t changed_capacity; /* entries of the changed-code list (more: full pair-list rebuild) */
  int64_t instances;      /* independent cutting instances packed in this state (1: a single solver) */
  int64_t active_bricks;  /* bricks holding particles (entries of brick_list; 0: every brick) */
  int64_t active_nodes;   /* nodes some particle's stencil touches (entries of node_list; 0: all) */
  int64_t pair_mode;      /* pair search (solver.py:260-263): 0 Verlet partner lists (list, auto),
                             1 partner lists rebuilt from the cell buckets every step (hash),
                             2 every pair q > p tested (all; O(n^2)) */
  int64_t regular;        /* every particle at local cell coordinates 1/4 or 3/4 per axis, in the
                             cell slot of its position bits (4 bx + 2 by + bz): the transfer
                             kernels may use the separable stencil */
  int64_t maxwell;        /* Maxwell branch on (maxwell_tau > 0 and maxwell_fraction > 0,
                             solver.py:189-191): the stress is that of maxwell_at (material.py:
                             485-535) and every F update relaxes the branch */
  int64_t tools;          /* knives (tool_count; 0 or 1: one).  Knife k > 0 is a free knife with its
                             own friction history row and damage (extra_blade, contact.py:500-566)
                             and labels digit 3^(cut index + k); the tool arrays hold k rows */
  double dt, spacing, mass, volume, radius, blade_offset, mu, lam, yield_stress, flip;
  double sn, gk, mus, mud, gc, board_sn, threshold;
  double half[3], bevel, blade_mass, inertia_inv[3], board_half[2];
  double drive_k, drive_c, drive_limit, rot_k, rot_c;
  double pair_skin;
  double viscosity;       /* dynamic viscosity (Pa s); > 0 adds viscous_impulse (material.py:321-340) */
  double mu_maxwell;      /* maxwell_fraction * mu */
  double maxwell_relax;   /* 1 - exp(-dt / maxwell_tau) */
  int64_t mover_capacity; /* > 0: a particle that moves too far from its pair-list build position
                             (a mover) only refreshes its own list; at most this many movers
                             per step and displaced particles per full build, else every list is
                             rebuilt.  0: every list is rebuilt.  One solver in the list mode with
                             at least 2^17 buckets and list_lanes <= 16 only */
} bm_params;

typedef struct bm_state {
  /* particles */
  float *x, *v, *F, *Fp, *stress, *damage, *ref, *local, *impulses, *histories;
  int64_t *codes;          /* side codes (contact.py:996-1039) */
  int64_t *transfer_codes; /* codes used by the grid transfers: codes, or zeros for single field */
  int32_t *base_node;
  uint8_t *field_idx;
  /* grid field table */
  int64_t *field_code;
  int32_t *node_nf;
  float *grid_mass, *momentum, *grid_impulse, *old_v, *new_v, *remapped, *grid_vel;
  /* tool and scalars */
  float *tool_pos, *tool_vel, *tool_rot, *tool_omega, *mobility;
  float *reaction, *torque, *board, *work, *speed, *table;
  float *interval_impulse, *interval_angular;
  int32_t *errors, *topology_dirty;
  int32_t *fp_identity;    /* [1] 1 while every Fp is exactly I (no plastic flow, no host write) */
  int32_t *update_pending; /* [1] 1 while the F update of the last substep waits for the next
                            * P2G (bm_substep_deferred), 2 when also some particles register
                            * again before it; null disables the deferral */
  int32_t *update_done;    /* [n] 1 for a particle registered again while its update waits: its
                            * earlier field slots are in field_idx_prev; zero between substeps */
  /* batched instances (null for a single solver): the instance of each particle and its Lame
   * constants; tool_pos/vel/rot/omega, mobility, reaction, torque, board, work and the
   * interval sums are then per instance ([instances][...]) */
  int32_t *instance;       /* [n] */
  float *inst_mu, *inst_lam; /* [instances] */
  /* [instances][16][10] contact sums of the contact kernel (board 3, reaction 3, torque 3,
   * work), spread over 16 slots by workgroup so that the atomics of an instance's many
   * wavefronts do not all meet on one address; zero between steps (the step's finish adds
   * and clears them) */
  float *inst_sums;
  /* per-instance pair lists: each instance's cell list lives in a bucket region of its own
   * ([instances][2]: first bucket, bucket count, a power of two; the regions tile
   * [0, list_buckets)), carries its own build tag and is rebuilt on its own request (set by
   * the instance's particles; pair_dirty rebuilds every list) */
  int32_t *inst_buckets;
  int32_t *inst_pair_builds; /* [instances] */
  int32_t *inst_pair_dirty;  /* [instances] */
  /* pair contacts: ping-pong history tables [2][P] and the partner lists */
  int64_t *pair_keys;      /* [2][P], -1 empty */
  float *pair_history;     /* [2][P][3] */
  int32_t *pair_used;      /* [2][P] */
  int32_t *pair_count;     /* [2] */
  int32_t *pair_dirty;     /* [1] rebuild request */
  int32_t *pair_builds;    /* [1] rebuilds so far; tags the bucket lists of the current build */
  float *list_reach;       /* [1] */
  float *built_x;          /* [n][3] positions at the last rebuild */
  int64_t *built_codes;    /* [n] */
  int32_t *partners;       /* [n][lanes] */
  int32_t *partner_count;  /* [n], -1 = more than lanes partners */
  int64_t *list_heads;     /* [B]: (build tag << 32) | first particle of the bucket */
  int32_t *list_links;     /* [n]: next particle of the same bucket, -1 at the end */
  /* pair-list updates after label changes only: the particles relabelled since the last
   * list update, and the particles whose lists are rebuilt at the build positions */
  int32_t *changed;        /* [changed_capacity] */
  int32_t *changed_count;  /* [1] */
  int32_t *pair_mark;      /* [n]: == pair_epoch + 1 for a list to rebuild this step */
  int32_t *pair_epoch;     /* [1] */
  /* brick bins: bricks x-major, cells of a brick z-fastest; cell c holds the particles
   * perm[8c .. 8c+8) in reference order, -1 for unused slots (at most 8 per cell) */
  int32_t *perm;           /* [8 * bricks * cells per brick] */
  int32_t *brick_list;     /* [active_bricks] bricks with particles, in brick order (null: all) */
  int32_t *node_list;      /* [active_nodes] nodes of the particles' stencils, ascending (null: all) */
  float *p2g_partial;      /* [bricks][F][W][7] window sums of mass, momentum, impulse */
  float *remap_partial;    /* [bricks][F][W][3] window sums of the re-scattered momentum */
  float *rigid_partial;    /* [rigid_blocks][12] per-block contact sums */
  uint32_t *counters;      /* [4]: [0] float bits of the step's largest particle speed (contact
                            * kernel, atomicMax; zero between steps), [1 + t] the insertion
                            * count of pair table t while a pair kernel fills it (copied to
                            * pair_count when the step's pair work is done), [3] unused */
  /* held-blade grid constraint (blade_constraint.py), used when blade_hold is set */
  float *bc_side;          /* [n] face side of the previous step (0, +1, -1) */
  float *bc_normal;        /* [n][3] */
  float *bc_lever;         /* [n][3] */
  float *bc_rhs, *bc_bound, *bc_impulse, *bc_extrapolated; /* [n] */
  int32_t *bc_active;      /* [n] particles with a face contact this step */
  int32_t *bc_nodes;       /* [N] nodes touched by them */
  int32_t *bc_node_mark;   /* [N] 0 between steps */
  float *bc_grid;          /* [F][N][3] 0 between steps */
  float *bc_coupled;       /* [F][N][3] */
  int32_t *bc_counts;      /* [4] active particles, active nodes, needs_solve, batches */
  float *bc_sums;          /* [8] linear 3, angular 3, max residual of the run, last residual */
  /* [2][ceil(n / 64)]: bit (p & 63) of word p >> 6 of row r is set when friction history row
   * r of particle p may be nonzero.  The contact kernel writes a zero history only over a
   * set bit; host writes to the histories must set every bit.  Null: always write. */
  uint64_t *history_mask;
  /* [n][9] float64 Fv - I of the Maxwell branch (f_maxwell = F Fp^-1 Fv^-1), used when
   * maxwell is set */
  double *viscous;
  /* [F][N][3] viscous grid impulses of a substep; all zero between substeps (the node
   * kernel that applies them clears them), used when viscosity > 0 */
  float *viscous_impulse;
  /* knives k = 1 .. tools-1 (used when tools > 1): friction history [tools-1][n][3], damage
   * [tools-1][n] and per-block contact sums [tools-1][rigid_blocks][8] (reaction 3, torque 3,
   * work, pad) */
  float *extra_histories;
  float *extra_damage;
  float *extra_partial;
  /* Nodes with 2..4 fields, for the node kernels' exchange lanes: the transfer kernel appends
   * each window node it finds with node_nf >= 2 once per epoch; the contact kernel empties
   * the list and advances the epoch at the start of a substep.  Stale or repeated entries
   * are harmless (the node kernels re-read node_nf).  Null: no list, the node threads do
   * the exchange themselves. */
  int32_t *multi_nodes;    /* [N] */
  int32_t *multi_stamp;    /* [N] epoch of a node's last append */
  int32_t *multi_count;    /* [2] entries, epoch (starts at 1) */
  uint8_t *field_idx_prev; /* [n][8] field slots before a registration that ran while the
                            * particle's F update waited (marked in update_done) */
  /* Pair-list refresh (mover_capacity > 0).  movers: the particles that moved more than half
   * the skin from their build positions in this step's G2P; the next contact kernel gives each a
   * new build position and list.  refresh_mark: a particle's last refresh time (2 epoch + 3 of
   * the step that marked it; zeros at the start).  A particle refreshed since the last full
   * build is displaced: the cell lists (of that build) hold it at a stale position, and the
   * searches take it from `displaced`.  A pair belongs to the list of the particle refreshed
   * last (the smaller index on a tie).  refresh_state: [0] movers, [1] displaced particles,
   * [2] refresh steps so far, [3] time of the last full build.  All zeros at the start. */
  int32_t *movers;         /* [mover_capacity] */
  int32_t *displaced;      /* [mover_capacity] */
  int32_t *refresh_mark;   /* [n] */
  int32_t *refresh_state;  /* [4] */
} bm_state;

/* Library identification; returns the number of exported stage entry points. */
int bm_version(void);

/* Stress of the current F, Fp (material.py:94-124, compute_stress; solver.py:318-329). */
int bm_refresh_stress(const bm_params* p, const bm_state* s, void* stream);

/* Zero reaction, torque, board, speed (contact.py:1058-1068, clear_step). */
int bm_clear_step(const bm_params* p, const bm_state* s, void* stream);

/* Gravity, board and blade contact with cutting resistance; max speed; pair-list
 * staleness check (contact.py:332-497 rigid_and_speed, contact.py:716-735 check_pair_list;
 * solver.py:588-630). */
int bm_rigid_contacts(const bm_params* p, const bm_state* s, int32_t cut_bits, void* stream);

/* Particle-pair contact with friction history over a sort-based cell list
 * (contact.py:569-925; solver.py:670-786).  parity = step_count % 2 selects the tables. */
int bm_pair_contacts(const bm_params* p, const bm_state* s, int32_t parity, void* stream);

/* Multi-field topology: rebuild the node field table and the per-corner field slots when
 * topology_dirty is set (material.py:37-75, 621-653; solver.py:391-429). */
int bm_register(const bm_params* p, const bm_state* s, void* stream);

/* P2G of mass, momentum and stress + contact impulse into the field table
 * (material.py:127-152; clears the used fields first, material.py:656-669). */
int bm_p2g(const bm_params* p, const bm_state* s, void* stream);

/* Compatible-field exchange (Eq. 8) and grid velocities (material.py:252-318). */
int bm_couple_and_velocity(const bm_params* p, const bm_state* s, void* stream);

/* G2P with PIC/FLIP and TLMPM position update, MUSL re-scatter (material.py:343-401). */
int bm_g2p_remap(const bm_params* p, const bm_state* s, void* stream);

/* Eq. 8 exchange of the re-scattered momentum, then field velocity = p / m
 * (solver.py:494, material.py:155-209). */
int bm_couple_remapped(const bm_params* p, const bm_state* s, void* stream);

/* MUSL F update, optional Hencky von Mises return map, stress for the next step
 * (material.py:404-475). */
int bm_update_F(const bm_params* p, const bm_state* s, void* stream);

/* Whole material step: register, p2g, couple_and_velocity, g2p_remap, couple_remapped,
 * update_F (solver.py:334-535). */
int bm_material_step(const bm_params* p, const bm_state* s, void* stream);

/* Knife drive and knife/board contact, one thread (tool.py:146-259; solver.py:824-864).
 * targets[3*index], rotations[4*index] give the servo target of this substep. */
int bm_tool_step(const bm_params* p, const bm_state* s, const float* targets, const float* rotations,
                 int32_t index, void* stream);

/* External tool: accumulate the interval impulse, optional mobility prediction, reset
 * the per-step sums (tool.py:70-115). */
int bm_finish_external(const bm_params* p, const bm_state* s, void* stream);

/* Interval sums for a driven knife (tool.py:6-16). */
int bm_accumulate_reaction(const bm_params* p, const bm_state* s, void* stream);

/* Side labels behind the blade edge; digit = 3^cut_index (contact.py:996-1039).  The
 * relabelled particles are recorded in `changed`, and their new candidate pairs are added to
 * the partner lists (bm_mark_relabelled). */
int bm_assign(const bm_params* p, const bm_state* s, int64_t digit, int32_t cutting, void* stream);

/* The partner-list side of a label change: each particle recorded in `changed` adds the pairs
 * that its new code makes candidates to the lists (no counterpart in the reference, which
 * rebuilds every list on any code change, contact.py:716-735).  bm_assign and bm_substep run
 * it themselves; a caller that writes codes and `changed` directly runs it before the next
 * pair contacts. */
int bm_mark_relabelled(const bm_params* p, const bm_state* s, void* stream);

/* One full substep in the reference order (solver.py:920-940).  external selects
 * finish_external_step instead of tool.step. */
int bm_substep(const bm_params* p, const bm_state* s, int32_t cut_bits, int32_t cutting, int64_t digit,
               int32_t parity, int32_t accumulate, const float* targets, const float* rotations,
               int32_t index, void* stream);

/* As bm_substep (the same Solver._step_kernels, solver.py:920-940), but its F update
 * (update_F_plastic, material.py:445-475) is left to the P2G of the next bm_substep or
 * bm_substep_deferred on the same stream, which updates each particle and uses its new stress
 * at once.  Particles that register again before that P2G (those around relabelled particles;
 * all of them after a full re-registration) are updated within this call.  Nothing may read
 * F, Fp or the stress in between; a sequence ends with bm_substep, which completes its own
 * update. */
int bm_substep_deferred(const bm_params* p, const bm_state* s, int32_t cut_bits, int32_t cutting, int64_t digit,
                        int32_t parity, int32_t accumulate, const float* targets, const float* rotations,
                        int32_t index, void* stream);

/* Held-blade grid constraint (blade_constraint.py:13-324): face contacts of the particles
 * beside the blade's broad faces, projected-gradient solve of W lambda >= rhs, lambda >= 0 on
 * the grid velocity new_v, reaction added to the tool.  Runs between couple_and_velocity and
 * g2p_remap; a no-op without face contacts that move into the blade. */
int bm_blade_constraint(const bm_params* p, const bm_state* s, void* stream);
/* result[size - 1] = largest unresolved inward velocity so far (blade_constraint.py:208-210). */
int bm_store_residual(const bm_params* p, const bm_state* s, float* result, int32_t size, void* stream);

/* External-tool port kernels (tool.py:19-68, material.py:6-20). */
int bm_load_external_state(const bm_params* p, const bm_state* s, const float* packed49, void* stream);
int bm_record_pose(const bm_params* p, const bm_state* s, int32_t index, float* out, void* stream);
int bm_store_interval(const bm_params* p, const bm_state* s, float* out, void* stream);
int bm_validate_interval(const bm_params* p, const bm_state* s, void* stream);

/* Particles beside the blade body (contact.py:1042-1055); count is int32[1]. */
int bm_count_embedded(const bm_params* p, const bm_state* s, int32_t* count, void* stream);

/* Energy diagnostics (diagnostics.py:7-35): out[6] = kinetic, gravity, elastic, max speed,
 * max z, min det F; the caller initialises out. */
int bm_metrics(const bm_params* p, const bm_state* s, float* out, void* stream);

/* Hencky von Mises projection of every particle (material.py:434-442, test probe). */
int bm_plastic_project(const bm_params* p, const bm_state* s, void* stream);

/* Probes of the contact functions (contact.py:21-43, 46-120, 181-232), count lanes each.
 * blade_query: in local[3*count] (device) and half[3] (host) -> out[7*count] = closest 3,
 * normal 3, distance. */
int bm_probe_blade_query(const float* local, const float* half, float bevel, float* out, int32_t count,
                         void* stream);
/* friction: in[13*count] = normal 3, relative 3, jn, history 3, me, (pad 2); params
 * dt, mus, mud, gk, gc; out[6*count] = impulse 3, history 3. */
int bm_probe_friction(const float* in, float dt, float mus, float mud, float gk, float gc, float* out,
                      int32_t count, void* stream);
/* Test probe of the contact kernel's cutting resistance (resist_cut, contact.py:182-232;
 * the reference's test tests/test_cut_resistance.py:15-42 launches it the same way): lane i
 * is a particle x[i], v[i] with damage[i] (updated in place) under the knife edge moving at
 * tool_velocity[i].  params19 = impulse j (3), knife origin (3), rotation xyzw (4), angular
 * velocity (3), half sizes (3), contact radius, grid spacing, damage work threshold.  out:
 * the returned impulse per lane; work[0] accumulates the work done (atomically). */
int bm_probe_resist_cut(const float* x, const float* v, const float* tool_velocity, float* damage,
                        const float* params19, int32_t cut_bits, float* out, float* work, int32_t count, void* stream);
/* 3x3 SVD of A[9*count] -> U[9], s[3], V[9] per lane. */
int bm_probe_svd3(const float* A, float* U, float* s, float* V, int32_t count, void* stream);
/* Test probe: fixed-corotated first Piola stress of each F (Fp = I) as the F update computes
 * it (stress_at, material.py:94-112); svd_path[i] = 0 polar iteration, 1 SVD, -1 det F <= 0. */
int bm_probe_stress(const float* F, float mu, float lam, float* out, int32_t* svd_path, int32_t count,
                    void* stream);

/* ---------------------------------------------------------------- piece labels
 * topology.components (topology.py:17-57): connected components of the particles through
 * the (grid node, cut code) fields of their positive-weight stencil corners, compatible
 * fields of one node joined; labels ranked by the lowest particle index of each piece.
 * The caller provides the scratch buffers. */
typedef struct bm_labels {
  int64_t fields;             /* L: slots per node of the scratch field table (1..255) */
  const uint8_t *corner_mask; /* [n] bit a set: the stencil weight of corner a is > 0 */
  int64_t *field_code;        /* [L][N] scratch field table */
  int32_t *node_nf;           /* [N] */
  uint8_t *field_idx;         /* [n][8] */
  int32_t *parent;            /* [n + L N] union-find forest */
  int32_t *root;              /* [n] */
  int32_t *block_sum;         /* [ceil(n / 1024)] */
  int32_t *labels;            /* [n] out */
  int32_t *count;             /* [1] out: number of pieces */
  int32_t *errors;            /* [1] out: nonzero when a node needed more than L fields */
} bm_labels;

/* Batched instances: per-instance interval load (instances x 49 floats: position, quaternion
 * xyzw, velocity, angular velocity, mobility) and store (instances x 6 floats: interval impulse
 * and angular impulse).  No reference counterpart: the batched driver of force_convergence. */
int bm_load_external_batch(const bm_params* p, const bm_state* s, const float* packed, void* stream);
int bm_store_interval_batch(const bm_params* p, const bm_state* s, float* out, void* stream);

/* Replaces the kernel couple (material.py:155-209) on caller tables: keys [count] (code *
 * node_count + node, -1 unused), heads [node_count] and links [count] (each node's fields as
 * a linked list, -1 ends), mass [count], state [count][3] -> result [count][3]; fields with a
 * negative key or no mass are not written.  The reference's test launches it on its own
 * tables (tests/test_mechanics.py:80-105). */
int bm_couple(const int64_t* keys, const int32_t* heads, const int32_t* links, const float* mass, const float* state,
              float* result, int64_t count, int64_t node_count, void* stream);
/* Replaces the kernel g2p (material.py:343-384) on caller tables: particles x, v [n][3]
 * (updated in place), slots and weights [n][8] into the grid arrays old_v, new_v
 * [grid_count][3] (tests/test_particle_transport.py:10-45 launches it this way). */
int bm_g2p(float* x, float* v, const int32_t* slots, const float* weights, const float* old_v, const float* new_v,
           float dt, float flip, int64_t n, int64_t grid_count, void* stream);

/* Replaces topology.components (topology.py:17-57) for a solver's current codes. */
int bm_piece_labels(const bm_params* p, const bm_state* s, const bm_labels* l, void* stream);

/* No counterpart in the reference.  bm_piece_labels without allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops): the
 * scratch buffers (field table of `fields` slots per node, union-find forest, roots, block
 * sums) live in a static device array of BM_LABELS_SCRATCH_BYTES on each GPU, kept until the
 * process ends.  The caller passes the inputs (solver codes and stencil through p and s, the
 * positive-weight corner mask [n]) and the outputs labels [n], count [1] (pieces) and
 * errors [1] (nonzero when a node needed more than `fields` codes: call again with more).
 * Returns -3 when bm_piece_labels_static_bytes(n, N, fields) exceeds the array, -4 when the
 * array's device address cannot be looked up.  Calls on one GPU share its array and keep
 * the last call's contents; calls that overlap on one GPU (two streams) race. */
#define BM_LABELS_SCRATCH_BYTES (64ll << 20)
int64_t bm_piece_labels_static_bytes(int64_t n, int64_t node_count, int64_t fields);
int bm_piece_labels_static(const bm_params* p, const bm_state* s, int64_t fields, const uint8_t* corner_mask,
                           int32_t* labels, int32_t* count, int32_t* errors, void* stream);

/* ---------------------------------------------------------------------------------
 * MLS-MPM cutting baseline (blademaster_warp/mlsmpm.py:42-229, MlsSolver._substep
 * mlsmpm.py:293-393).  Dense grid of dims nodes with planes [3][nodes] for momentum and
 * velocity; grid_m and grid_p must be zero before the first P2G (the grid update leaves
 * them zero for the next one). */
typedef struct bm_mls_params {
  int64_t n;
  int64_t dims[3];
  int64_t plastic;         /* yield_stress > 0 */
  double origin[3];
  double dx, dt, mass, volume, mu, lam, damaged_shear, softening, compressed, stretched;
  int64_t groups;          /* workgroups of P2G when perm is set (kBlock = 256 slots each) */
} bm_mls_params;

typedef struct bm_mls_state {
  float *x, *v, *C, *F, *damage, *damage_reference, *yield_stress, *impulse;
  float *grid_m, *grid_p, *grid_v;
  float *interval_impulse, *interval_angular;
  int32_t *inversions;
  /* [groups][256] particle of each P2G slot, -1 for none: particles grouped by bricks of
   * 3 x 3 x 3 stencil base cells, so a workgroup's particles share nodes.  Null: P2G takes
   * particles in index order.  Any grouping gives the same sums up to float order. */
  int32_t *perm;
} bm_mls_state;

/* Quadratic B-spline APIC P2G with the corotated stress (mlsmpm.py:75-119). */
int bm_mls_p2g(const bm_mls_params* m, const bm_mls_state* s, void* stream);
/* Nodal velocity with sticky walls (mlsmpm.py:122-137); zeroes grid_m and grid_p. */
int bm_mls_grid_update(const bm_mls_params* m, const bm_mls_state* s, void* stream);
/* APIC G2P, F update with singular-value clamp, return map, damage (mlsmpm.py:157-229). */
int bm_mls_g2p(const bm_mls_params* m, const bm_mls_state* s, void* stream);
/* One MLS substep: knife/board contact on the MLS particles (p, t as for bm_rigid_contacts
 * with tool_model -1), interval sums, P2G, grid update, G2P (mlsmpm.py:293-393). */
int bm_mls_substep(const bm_params* p, const bm_state* t, const bm_mls_params* m, const bm_mls_state* s,
                   void* stream);

/* ---------------------------------------------------------------------------------
 * Interval-wrench surrogate.  No reference counterpart: the reference's only intra-interval
 * predictor is the mobility matrix of tool_substep "predict" (tool.py:97-111); this network
 * predicts the knife's interval wrench from the knife pose and twist, the previous interval's
 * wrench and a pooled occupancy/damage patch around the edge, for couplings that exchange
 * less often than once per robot step.
 *
 * Network: inputs -> hidden (ReLU) -> hidden (ReLU) -> outputs.  Layer l is an augmented
 * matrix Wa_l [rows_l + 1][cols_l], row-major, whose last row is the bias, so a layer is
 * [x, 1] Wa.  params = Wa_1 | Wa_2 | Wa_3 (float32); grads, adam_m, adam_v have its layout.
 * Training is float32 (MFMA 16x16x4 f32); inference is bfloat16 weights and activations with
 * float32 accumulation (MFMA 16x16x32 bf16) from a packed copy of params. */
typedef struct bm_mlp {
  int32_t inputs, hidden, outputs;
  float *params, *grads, *adam_m, *adam_v;
  float *h1, *h2, *out;          /* activations [batch][hidden], [batch][hidden], [batch][outputs] */
  float *d_out, *d_h2, *d_h1;    /* their gradients */
  float *loss;                   /* [1] mean squared error of the last forward/backward */
  int32_t *step;                 /* [1] Adam steps taken; incremented by every training step */
  const float *lr;               /* [2] Adam learning rate, decoupled weight decay (AdamW) */
  uint16_t *packed;              /* bf16 inference weights; bm_mlp_sizes gives the length */
  float *work;                   /* split-K partial products of the weight gradients (may be null) */
  int64_t work_len;              /* floats in work; bm_mlp_sizes gives the largest useful length */
  /* Standardisation applied by bm_mlp_infer only (each null: identity), in float32:
   * x' = (x - x_shift) x_scale per input, y = y' y_scale + y_shift per output. */
  const float *x_shift, *x_scale, *y_shift, *y_scale;
} bm_mlp;

/* Lengths of params, of packed and of the largest useful work buffer for a network shape. */
int bm_mlp_sizes(int32_t inputs, int32_t hidden, int32_t outputs, int64_t* params, int64_t* packed,
                 int64_t* work);
/* out = net(x), float32; x [batch][inputs]. */
int bm_mlp_forward(const bm_mlp* m, const float* x, int32_t batch, void* stream);
/* Forward, mean squared error against target [batch][outputs], and grads; params unchanged. */
int bm_mlp_backward(const bm_mlp* m, const float* x, const float* target, int32_t batch, void* stream);
/* bm_mlp_backward followed by one AdamW step (beta1 0.9, beta2 0.999, eps 1e-8). */
int bm_mlp_train_step(const bm_mlp* m, const float* x, const float* target, int32_t batch, void* stream);
/* packed = bf16 copy of params (transposed, zero-padded); needs hidden 256, inputs < 128,
 * outputs <= 16. */
int bm_mlp_pack(const bm_mlp* m, void* stream);
/* y [batch][outputs] = net(x) from packed, one fused launch. */
int bm_mlp_infer(const bm_mlp* m, const float* x, float* y, int32_t batch, void* stream);
/* Raw GEMM of the training path, for tests: C [M][N] = op(A) op(B), op(A) [M][K], op(B) [K][N];
 * trans_a: A is stored [K][M]; trans_b: B is stored [N][K]. */
int bm_mlp_gemm(const float* A, const float* B, float* C, int32_t M, int32_t N, int32_t K, int32_t lda,
                int32_t ldb, int32_t ldc, int32_t trans_a, int32_t trans_b, void* stream);
/* The same product with K split over workgroups (partials in work, then a fixed-order sum);
 * ldc = N.  This is how the weight gradients are formed when the tile grid is small. */
int bm_mlp_gemm_split(const float* A, const float* B, float* C, int32_t M, int32_t N, int32_t K, int32_t lda,
                      int32_t ldb, int32_t trans_a, int32_t trans_b, float* work, int64_t work_len, void* stream);

/* Occupancy/damage patch around each knife edge (input of the surrogate).  The box spans
 * [-4h, 4h] across the blade (local x), the blade's half length along the edge (local y) and
 * [-4h, 4h] about the edge (local z), in 8 x 8 x 8 cells; per instance and cell, out holds the
 * particle count and the summed damage: out [instances][2][8][8][8].  Cleared first. */
int bm_blade_patch(const bm_params* p, const bm_state* s, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BM_HIP_H */
