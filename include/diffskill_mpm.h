/*
 * diffskill_mpm.h -- C ABI of the MI355X (gfx950) differentiable MLS-MPM engine.
 *
 * Plain pointers and sizes only.  All device pointers are HIP device memory;
 * every call that launches work takes the hipStream_t (as void*) to run on.
 *
 * Each entry point replaces a piece of the reference's Taichi path
 * (Xingyu-Lin/DiffSkill, plb/engine):
 *
 *   mpm_create / mpm_workspace_bytes   MPMSimulator.__init__ field allocation
 *                                      (mpm_simulator.py:8-67) and
 *                                      Primitives(...) (primitives.py:821-838)
 *   mpm_set_frame / mpm_get_frame      setframe / readframe, set_state /
 *                                      get_state (mpm_simulator.py:348-396)
 *   mpm_reset_frame0                   reset_kernel (mpm_simulator.py:402-413)
 *   mpm_set_prim_frame / mpm_get_prim_frame
 *                                      Primitive.set_state / get_state
 *                                      (primive_base.py:182-207)
 *   mpm_set_actions                    Primitive.set_action + set_velocity
 *                                      (primive_base.py:260-274)
 *   mpm_substep_fwd                    MPMSimulator.substep (mpm_simulator.py:307-323)
 *   mpm_substep_bwd                    MPMSimulator.substep_grad (mpm_simulator.py:325-345)
 *   mpm_grid_mass / mpm_grid_mass_bwd  compute_grid_m_kernel and its .grad
 *                                      (mpm_simulator.py:456-466, function.py:147-152)
 *   mpm_density_loss_bwd               compute_density_loss_kernel.grad followed by
 *                                      compute_grid_m_kernel.grad (losses/loss.py:173-175,
 *                                      :213, :290), one launch
 *   mpm_add_obs_grad                   GradModel._set_obs_grad (function.py:130-142)
 *   mpm_action_grad                    set_velocity.grad + get_action_grad
 *                                      (function.py:180-190, primive_base.py:276-282)
 *   mpm_clear_grads                    GradModel.reset grad clearing (function.py:44-61)
 *   mpm_collision_bytes / mpm_set_collisions / mpm_get_collisions
 *                                      set_surface_points / set_collision_idx /
 *                                      apply_collision_projection and their .grad
 *                                      (mpm_simulator.py:286-305,316-320,337-339)
 *   mpm_last_error                     Python exceptions of the reference
 *
 * Threading: one context per stream; no internal locking (the reference is
 * single-threaded per process).  Every function returns 0 on success and a
 * negative code on error; mpm_last_error() describes the last error of the
 * calling thread.
 */
#ifndef DISKILL_MPM_H
#define DISKILL_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Primitive shape codes (class names of plb/engine/primitive/primitives.py). */
enum {
    MPM_SPHERE = 0, MPM_CAPSULE = 1, MPM_ROLLINGPIN = 2, MPM_ROLLINGPIN_EXT = 3,
    MPM_BOX = 4, MPM_GRIPPER = 5, MPM_PRISM = 6, MPM_KNIFE = 7,
    MPM_CHOPSTICKS = 8, MPM_GRIPPER2 = 9, MPM_CYLINDER = 10, MPM_TORUS = 11
};

/* One scene: the SIMULATOR constants of MPMSimulator.__init__
 * (mpm_simulator.py:8-39), derived on the host in double precision. */
typedef struct {
    int n_grid;            /* nodes per axis */
    int n_particles;       /* particles of this scene */
    int n_prims;           /* primitives of this scene */
    int pad0;
    double dx, inv_dx, dt, p_vol, p_mass;
    double mu, lam, yield_stress;
    double gravity[3];
    double ground_friction;
    double lower_bound;
} mpm_scene_desc;

/* One primitive (Primitive cfg + per-class parameters). */
typedef struct {
    int type;              /* MPM_* code */
    int action_dim;
    double friction, softness;
    double size[3];        /* Box/Gripper half extents; Knife box size */
    double h, r;           /* Capsule family h (segment length), r; Sphere radius in r;
                              Cylinder h (disk radius), r (half height); Torus tx in h, ty in r */
    double ph[2];          /* Prism / Knife prism h */
    double prot[4];        /* Prism / Knife prism rotation (w, x, y, z) */
    double lo[3], hi[3];   /* xyz_limit */
    double min_gap, max_gap;
    double action_scale[7];
} mpm_prim_desc;

typedef struct mpm_ctx mpm_ctx;

/* Bytes of device workspace a context needs: particle frames (checkpoint
 * every ckpt_k frames plus one live segment of ckpt_k - 1 frames; ckpt_k >=
 * max_frames means store-all), two adjoint frames, grids, primitive frames.
 * keep_blocks sets how many active 4x4x4 node blocks of each substep's grid
 * (and that substep's binning) are kept per frame slot for the adjoint, which
 * then skips re-binning and rebuilding the grid: < 0 picks a default from the
 * particle count, 0 keeps nothing (the adjoint rebuilds every grid), > 0 is
 * the capacity.  A substep whose active list exceeds it is rebuilt. */
int64_t mpm_workspace_bytes(const mpm_scene_desc *scenes, int n_scenes, int max_frames, int ckpt_k,
                            int keep_blocks);

/* Create a context over a caller-owned workspace (device pointer, >=
 * mpm_workspace_bytes).  prims: concatenated primitives of all scenes, in
 * scene order.  The workspace must stay alive until mpm_destroy. */
int mpm_create(const mpm_scene_desc *scenes, int n_scenes, const mpm_prim_desc *prims,
               int max_frames, int ckpt_k, int keep_blocks, void *workspace,
               int64_t workspace_bytes, void *stream, mpm_ctx **out);
int mpm_destroy(mpm_ctx *ctx);

/* Device pointer of frame f's particle state (SoA, 24 rows of `stride`
 * floats: x0 x1 x2 v0 v1 v2 C00..C22 F00..F22; scene k owns columns
 * [particle_offset(k), +n_particles)).  Recomputes the frame from its
 * checkpoint when it is not resident (forward substeps on `stream`). */
int mpm_frame(mpm_ctx *ctx, int f, void *stream, float **ptr);
int mpm_stride(mpm_ctx *ctx);
int mpm_particle_offset(mpm_ctx *ctx, int scene);
/* Adjoint of frame f (valid between backward substeps, same layout). */
int mpm_adjoint(mpm_ctx *ctx, int f, float **ptr);
/* Primitive frame state [max_frames][n_prims_total][8] (pos, quat wxyz,
 * gap), its adjoint, per-substep velocity [max_frames][n][8] (v, w,
 * gap_vel), its adjoint. */
int mpm_prim_buffers(mpm_ctx *ctx, float **state, float **state_grad, float **vel, float **vel_grad);

/* Copy a whole frame (all scenes, SoA layout above) from/to device memory. */
int mpm_set_frame(mpm_ctx *ctx, int f, const float *src, void *stream);
int mpm_get_frame(mpm_ctx *ctx, int f, float *dst, void *stream);
/* reset_kernel: frame 0 from positions x (3 rows of `stride` floats, the
 * SoA layout above), v = 0, F = I, C = 0. */
int mpm_reset_frame0(mpm_ctx *ctx, const float *x, void *stream);
/* Primitive state of frame f for every primitive ([n_prims_total][8], device). */
int mpm_set_prim_frame(mpm_ctx *ctx, int f, const float *src, void *stream);
int mpm_get_prim_frame(mpm_ctx *ctx, int f, float *dst, void *stream);

/* Set the actions of env step s for every primitive (device
 * [n_prims_total][7], already clipped), filling the per-substep velocity of
 * substeps [s*n_substeps, (s+1)*n_substeps). */
int mpm_set_actions(mpm_ctx *ctx, int s, int n_substeps, const float *actions, void *stream);

/* `count` forward substeps starting at frame s (all scenes). */
int mpm_substep_fwd(mpm_ctx *ctx, int s, int count, void *stream);
/* `count` adjoint substeps: frames s+count-1 down to s (LIFO).  Returns an
 * error when the adjoint of frame s+count is not the current one. */
int mpm_substep_bwd(mpm_ctx *ctx, int s, int count, void *stream);

/* Particle mass grid of frame f for every scene, ADDED into out
 * [n_scenes][n^3] (i, j, k row-major, node I at world position I*dx). */
int mpm_grid_mass(mpm_ctx *ctx, int f, float *out, void *stream);
/* Adds d(sum g*m)/dx into the adjoint of frame f. g: [n_scenes][n^3]. */
int mpm_grid_mass_bwd(mpm_ctx *ctx, int f, const float *g, void *stream);
/* Adjoint of the weighted L1 density loss weight * sum |m - target| at frame
 * f: adds d/dx into the adjoint of frame f, with the node adjoint
 * weight * sign(m - target) (0 where they are equal) formed inside the
 * grid-mass adjoint.  mass: the mass grid of frame f (mpm_grid_mass on a
 * zeroed grid), target: the target densities, both [n_scenes][n^3]. */
int mpm_density_loss_bwd(mpm_ctx *ctx, int f, const float *mass, const float *target, float weight, void *stream);
/* Adds observation adjoints at frame f: xv_grad [stride][6] (x, v of every
 * particle slot; may be NULL), prim_grad [n_prims_total][8] (may be NULL). */
int mpm_add_obs_grad(mpm_ctx *ctx, int f, const float *xv_grad, const float *prim_grad, void *stream);
/* Action adjoint of env step s: out [n_prims_total][7]. */
int mpm_action_grad(mpm_ctx *ctx, int s, int n_substeps, float *out, void *stream);
/* Zero every adjoint (particles and primitives) and make frame f the top
 * of the adjoint chain; f = -1 lets the first adjoint call choose it. */
int mpm_clear_grads(mpm_ctx *ctx, int f, void *stream);
/* Re-upload the primitive table (softness / friction / limits; types fixed).
 * Replaces Primitives.set_softness (primitives.py:876-878). Synchronizes `stream`. */
int mpm_update_prims(mpm_ctx *ctx, const mpm_prim_desc *prims, void *stream);
/* Copy frame src (particles and primitive state) to frame dst and restart
 * the trajectory there: copyframe (mpm_simulator.py:369-378), used by
 * MPMSimulator.step in copy mode (mpm_simulator.py:449-451). */
int mpm_copy_frame(mpm_ctx *ctx, int src, int dst, void *stream);

/* compute_min_dist (plb/engine/function.py:79-88): signed distance of every
 * particle of frame f to every primitive of its scene -- Primitive.sdf, or
 * for Gripper / Gripper2 both jaws (sdf_2 with flag -1, then +1) -- in
 * primitive order, into out[col][stride] (device, n_cols rows; columns past a
 * scene's own count are left untouched).  mpm_min_dist_cols gives the column
 * count of one scene, or the largest over all scenes for scene < 0. */
int mpm_min_dist_cols(mpm_ctx *ctx, int scene);
int mpm_min_dist(mpm_ctx *ctx, int f, float *out, int n_cols, void *stream);
/* compute_min_dist.grad (function.py:154-160): gout[col][stride] adjoints of
 * the distances; adds their d/dx into the adjoint of frame f (which must be
 * the current adjoint frame) and the primitive-state adjoint of frame f. */
int mpm_min_dist_bwd(mpm_ctx *ctx, int f, const float *gout, int n_cols, void *stream);

/* Tool-tool collision projection (MPMSimulator.set_num_collision,
 * initialize and substep, mpm_simulator.py:69-96,286-320).  Every forward
 * substep, after forward kinematics, each pair (i, j) samples n_points
 * points of box j (rand_num[k] * size_j clamped to the box, moved by j's
 * pose), takes the point with the smallest negative SDF of primitive i (the
 * first on ties; none if no point is inside) and moves i by
 * normalized(normal_i) * sdf_i at that point; all points and choices use the
 * poses after kinematics, and the moves apply in pair order.  The adjoint
 * (substep_grad, mpm_simulator.py:337-339) reaches i's position, rotation
 * and gap; the points carry none (rand_points has no gradient).
 *
 * mpm_collision_bytes: device bytes the pairs' tables and per-frame records
 * need (-1 on bad sizes).  mpm_set_collisions: pairs[2p], pairs[2p+1] = (i, j)
 * as primitive indices of the context, both of one scene, grouped by scene
 * in ascending order and in the reference's order within a scene; j must be
 * a Box or Gripper (Box._project, primitives.py:369-372); rand_num is a host
 * array n_pairs x n_points x 3 (the reference draws it with
 * np.random.seed(42); uniform(-1.5, 1.5)); buffer is caller-owned device
 * memory of at least mpm_collision_bytes that must outlive the context or
 * the next call.  n_pairs = 0 switches the pass off.  Synchronizes stream.
 * mpm_get_collisions: frame f's records (f in [1, max_frames), written by
 * the substep f-1 -> f): idx[n_pairs] collision_idx (-1: no contact),
 * points[n_pairs][3] the chosen points, pre[n_pairs][3] the position of i
 * before the pair's move (host arrays; any may be NULL).  Synchronizes. */
int64_t mpm_collision_bytes(mpm_ctx *ctx, int n_pairs, int n_points);
int mpm_set_collisions(mpm_ctx *ctx, int n_pairs, const int *pairs, int n_points, const float *rand_num, void *buffer,
                       int64_t buffer_bytes, void *stream);
int mpm_get_collisions(mpm_ctx *ctx, int f, int *idx, float *points, float *pre, void *stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every engine kernel (no counterpart in the reference; measurement only).
 * Phases (MPM_PH_COMPACT is kept for numbering and no longer launched: the
 * occupied and active block lists are built inside the bin/blocks kernels): */
enum {
    MPM_PH_BIN = 0, MPM_PH_COMPACT = 1, MPM_PH_BLOCKS = 2, MPM_PH_SCATTER = 3, MPM_PH_STRESS = 4,
    MPM_PH_P2G = 5, MPM_PH_GRID = 6, MPM_PH_G2P = 7, MPM_PH_P2G_GRID = 8, MPM_PH_GRID_KEEP = 9,
    MPM_PH_G2P_BWD = 10, MPM_PH_GRID_BWD = 11, MPM_PH_P2G_BWD = 12, MPM_PH_STRESS_BWD = 13,
    MPM_PH_COLLIDE = 14, MPM_PH_COLLIDE_BWD = 15, MPM_PH_PRIM_REDUCE = 16,
    MPM_N_PHASES = 17
};
int mpm_set_profiling(mpm_ctx *ctx, int on);

/* Launch shape of the particle and grid phases (no counterpart in the
 * reference, whose launches Taichi shapes): particles per work item (a
 * multiple of 256 in [256, 8192]; default 1024), persistent workgroups of the
 * work-item kernels and of the grid kernels (defaults 4 and 8 per CU).  0
 * keeps a value.  Results do not depend on it beyond float rounding; it is
 * what the autotuner varies.  Call outside graph capture. */
int mpm_set_launch(mpm_ctx *ctx, int particles_per_item, int tile_ctas, int grid_ctas);
int mpm_get_launch(mpm_ctx *ctx, int *particles_per_item, int *tile_ctas, int *grid_ctas);
/* Accumulated milliseconds and launch counts per phase (synchronizes the
 * recorded events); reset != 0 clears the accumulators. */
int mpm_profile_read(mpm_ctx *ctx, double *ms, long long *counts, int reset);

/* Fixed-order execution (no counterpart in the reference, whose Taichi
 * kernels accumulate with unordered atomics): with on != 0 every sum the
 * engine forms is taken in an order fixed by the inputs alone -- particles
 * ranked within their cell in slot order, each wave's tile contributions
 * summed in LDS tiles of its own and the tiles in wave order, the grid
 * adjoint's primitive adjoints per node block summed in block order, the
 * distance adjoint per particle wave in wave order, the mass grid
 * accumulated in integers (2^-40 p_mass) -- so identical inputs and launch
 * shape (mpm_set_launch's particles per work item) give bitwise identical
 * frames, mass grids and adjoints, and a checkpointed adjoint equals the
 * store-all one.  On by default; on = 0 selects unordered atomics for the
 * cell ranks and those three sums (about 1 % faster on a 64-scene batch,
 * results equal within float rounding, not bitwise from run to run).
 * Call outside graph capture. */
int mpm_set_deterministic(mpm_ctx *ctx, int on);
int mpm_get_deterministic(mpm_ctx *ctx);

/* Slot re-sort (no counterpart in the reference, whose particle fields have
 * one fixed order): with substeps = R >= 8, every frame s with s % R == 0
 * (0 < s < max_frames - 1) is written with its particles' slots re-sorted
 * into the cell order of substep s-1's binning, within each scene's columns,
 * so the tile kernels' gathers and scatters keep touching whole cache lines
 * as the material deforms; the adjoint maps back across each boundary, a
 * checkpointed recompute applies the orders the first pass defined, and the
 * results are those of the unsorted run up to float summation order (bitwise
 * reproducible in the fixed-order mode).  0 (the default) turns it off.
 * The caller passes a device buffer of at least mpm_resort_bytes(ctx, R)
 * bytes (the boundaries' permutations, a staging frame, per-block counts),
 * kept until the re-sort is turned off or the context destroyed.  Setting it
 * invalidates every frame after 0.  mpm_frame and the other per-particle
 * accessors of frame f then see frame f's own slot order; mpm_slot_order
 * writes, for every slot c of frame 0, the slot its particle holds in frame f
 * (stride ints, device memory; synchronises the stream).  mpm_copy_frame
 * refuses frames of different slot orders. */
int64_t mpm_resort_bytes(mpm_ctx *ctx, int substeps);
int mpm_set_resort(mpm_ctx *ctx, int substeps, void *buffer, int64_t buffer_bytes);
int mpm_get_resort(mpm_ctx *ctx);
int mpm_slot_order(mpm_ctx *ctx, int f, int *map, void *stream);

const char *mpm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
