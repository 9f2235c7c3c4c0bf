/*
 * vbd_hip.h -- C ABI of libvbd_hip.so, the MI355X (gfx950) implementation of
 * the particle Vertex Block Descent solver step of Newton's SolverVBD.
 *
 * Every pointer named in the descriptors is a DEVICE pointer unless its name
 * ends in _host.  Arrays use the reference's layouts: vec3 = 3 contiguous
 * float32 (AoS), mat22/mat33 row-major, indices int32.  No call allocates
 * device memory after vbd_create / vbd_collide_create, none but
 * vbd_iterate_to_residual synchronises the stream, and every kernel is enqueued
 * on the stream passed in, so a whole substep can be captured into a HIP graph.  Functions return 0 on success or
 * a negative VBD_ERR_* code; vbd_last_error() returns the message of the
 * calling thread's last failure.  Nothing throws across this ABI.
 *
 * Reference interfaces replaced (CSYSLAM/newton, paths relative to the
 * repository root):
 *   SolverVBD.__init__ particle part  newton/_src/solvers/vbd/solver_vbd.py:221-618 -> vbd_create
 *   SolverVBD.step                    newton/_src/solvers/vbd/solver_vbd.py:1773-1825 -> vbd_step
 *   _initialize_particles/forward_step solver_vbd.py:2065-2100, particle_vbd_kernels.py:1350-1375 -> vbd_forward
 *   init_body_particle_contacts        rigid_vbd_kernels.py:3276-3321 (via solver_vbd.py:2493-2514) -> vbd_contact_prepare
 *   update_duals_body_particle_contacts rigid_vbd_kernels.py:4861-4926 -> vbd_iteration
 *   _solve_particle_iteration (one color: accumulate_particle_body_contact_force_and_hessian
 *     + solve_elasticity(_tile) + apply_truncation_ts) solver_vbd.py:2552-2702 -> vbd_color_pass
 *   _finalize_particles/update_velocity solver_vbd.py:3158-3169, particle_vbd_kernels.py:1468-1471 -> vbd_finalize
 *   create_soft_contacts (particle vs plane/sphere) newton/_src/sim/collide.py:1553-1592,
 *     newton/_src/geometry/kernels.py:1020-1184 -> vbd_collide_generate / vbd_collide_generate_bodies
 *   launch_soft_ef_contacts (full-surface edge/face records) collide.py:1593-1610,
 *     newton/_src/geometry/soft_contacts_sdf.py:240-551 -> the same calls (edge/face pairs in the desc)
 *   SolverBase.integrate_particles     newton/_src/solvers/solver.py:20-56, :297-330 -> vbd_integrate_particles
 */
#ifndef VBD_HIP_H
#define VBD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBD_OK 0
#define VBD_ERR_ARG -1
#define VBD_ERR_HIP -2
#define VBD_ERR_STATE -3

typedef struct VbdContext VbdContext;
typedef struct VbdCollide VbdCollide;

/* Static model arrays (Model fields of newton/_src/sim/model.py:745-1019). */
typedef struct VbdModelDesc {
    int32_t particle_count;
    int32_t tri_count;
    int32_t edge_count;
    int32_t tet_count;
    const float* particle_mass;
    const float* particle_inv_mass;
    const float* particle_radius;
    const int32_t* particle_flags;
    const int32_t* tri_indices;            /* [T,3] */
    const float* tri_poses;                /* [T,2,2] */
    const float* tri_materials;            /* [T,5] */
    const float* tri_areas;                /* [T] */
    const int32_t* edge_indices;           /* [E,4] (o0,o1,v0,v1) */
    const float* edge_rest_angle;          /* [E] */
    const float* edge_rest_length;         /* [E] */
    const float* edge_bending_properties;  /* [E,2] */
    const int32_t* tet_indices;            /* [K,4] */
    const float* tet_poses;                /* [K,3,3] */
    const float* tet_materials;            /* [K,3] */
    const int32_t* v_adj_tris;             /* CSR of (element, slot) pairs */
    const int32_t* v_adj_tris_offsets;     /* [N+1] (counts ints, 2 per entry) */
    const int32_t* v_adj_edges;
    const int32_t* v_adj_edges_offsets;
    const int32_t* v_adj_tets;
    const int32_t* v_adj_tets_offsets;
    int32_t color_count;
    const int32_t* color_offsets_host;     /* [C+1] host */
    const int32_t* color_particles;        /* concatenated color groups (every particle once) */
    const float* gravity;                  /* vec3; world 0 is used for every particle */
    int32_t shape_count;
    const float* shape_margin;             /* [S] */
    const float* shape_material_ke;
    const float* shape_material_kd;
    const float* shape_material_mu;
    float soft_contact_ke;
    float soft_contact_kd;
    float soft_contact_mu;
    /* Model.spring_* (newton/_src/sim/model.py); springs are not part of the coloring graph, so both
     * endpoints of one spring may share a color */
    int32_t spring_count;
    const int32_t* spring_indices;         /* [S*2] */
    const float* spring_rest_length;       /* [S] */
    const float* spring_stiffness;         /* [S] */
    const float* spring_damping;           /* [S] */
    /* Model.shape_body ([shape_count]; NULL: every shape static).  Shapes attached to bodies move with
     * poses an external rigid solver gives each step (SolverVBD(integrate_with_external_rigid_solver=True),
     * solver_vbd.py:280-281, :2523-2526), handed over in VbdContactsDesc.body_q / body_q_prev */
    const int32_t* shape_body;
    int32_t body_count;
} VbdModelDesc;

typedef struct VbdSolverParams {
    int32_t iterations;
    float friction_epsilon;
    float avbd_beta;            /* rigid_avbd_linear_beta (0: penalty fixed at the averaged ke) */
    float k_start;              /* < 0: penalty seeded at the averaged ke */
    int32_t fp64_ill_conditioned; /* 1: blocks with Frobenius condition > 1e6 are re-solved in FP64 */
    int32_t soft_contact_capacity; /* largest soft_contact_max the solver will be given */
    /* cloth self-contact: SolverVBD(particle_enable_self_contact=True, particle_*) (solver_vbd.py:231-242,
     * :556-606); all pointers are device arrays, filters may be null */
    int32_t self_contact;                   /* 0: off */
    float self_contact_radius;              /* particle_self_contact_radius */
    float self_contact_margin;              /* particle_self_contact_margin (detection query radius) */
    float conservative_bound_relaxation;    /* particle_conservative_bound_relaxation */
    float max_displacement;                 /* float32(margin * relaxation * 0.5) evaluated in double */
    int32_t vertex_contact_buffer_size;     /* triangles kept per vertex (<= 64) */
    int32_t edge_contact_buffer_size;       /* edges kept per edge (<= 64) */
    int32_t collision_detection_interval;   /* < 0: before initialization only; 0: also at iteration 0; n: every n */
    float edge_parallel_epsilon;
    float rest_shape_contact_exclusion_radius;
    const float* particle_q_rest;           /* [N,3] rest positions (model.particle_q at creation) */
    const int32_t* particle_world;          /* [N], null: one world */
    const int32_t* vt_filter;               /* sorted triangle ids per vertex */
    const int32_t* vt_filter_offsets;       /* [N+1] */
    const int32_t* ee_filter;               /* sorted edge ids per edge */
    const int32_t* ee_filter_offsets;       /* [E+1] */
    /* 1: the contacts may hold full-surface edge / face records (CollisionPipeline(
     * enable_rigid_soft_full_surface_contact=True), collide.py:785-793); each adds up to three corner
     * entries, so the per-corner buffers are sized 3 x soft_contact_capacity */
    int32_t soft_contact_full_surface;
} VbdSolverParams;

/* Contacts.soft_contact_* (newton/_src/sim/contacts.py:327-371) and the body poses of the step. */
typedef struct VbdContactsDesc {
    const int32_t* soft_contact_count;  /* [1] device */
    int32_t soft_contact_max;
    const int32_t* soft_contact_indices; /* [S,3] (p,-1,-1) particle, (v0,v1,-1) edge, (v0,v1,v2) face records */
    const int32_t* soft_contact_shape;
    const float* soft_contact_body_pos;  /* body frame of the shape's body (world for static shapes) */
    const float* soft_contact_body_vel;
    const float* soft_contact_normal;    /* world */
    const float* soft_contact_barycentric; /* [S,3] corner weights; NULL: edge/face records are skipped */
    /* poses of the shapes' bodies ([body_count,7], p then q xyzw): body_q at the step's end (the external
     * rigid solver's state_out.body_q), body_q_prev at its start (state_in.body_q); NULL for static shapes */
    const float* body_q;
    const float* body_q_prev;
    float dt;                            /* the step's dt (velocity of body points); vbd_step sets it itself */
} VbdContactsDesc;

/* Writable soft-contact outputs of the collision step. */
typedef struct VbdContactsOut {
    int32_t* soft_contact_count;        /* [1] */
    int32_t soft_contact_max;
    int32_t* soft_contact_particle;     /* [S] */
    int32_t* soft_contact_indices;      /* [S,3] */
    float* soft_contact_barycentric;    /* [S,3] */
    int32_t* soft_contact_shape;        /* [S] */
    float* soft_contact_body_pos;       /* [S,3] */
    float* soft_contact_body_vel;       /* [S,3] */
    float* soft_contact_normal;         /* [S,3] */
} VbdContactsOut;

/* Particle and shape data for soft-contact generation. */
typedef struct VbdCollideDesc {
    int32_t particle_count;
    int32_t pair_count;
    const int32_t* pairs;               /* [P,2] (particle, shape), world-compatible order */
    const float* particle_radius;
    const int32_t* particle_flags;
    const int32_t* particle_world;
    int32_t shape_count;
    const int32_t* shape_type;          /* 1 plane, 3 sphere */
    const int32_t* shape_flags;
    const int32_t* shape_world;
    const float* shape_transform;       /* [S,7] (p, q xyzw), relative to the shape's body */
    const float* shape_scale;           /* [S,3] */
    const float* shape_margin;          /* [S] */
    const int32_t* shape_body;          /* [S]; NULL: every shape static */
    /* full-surface passes (collide.py:1598-1610, soft_contacts_sdf.py:277-551); counts 0 when off */
    int32_t edge_pair_count;
    const int32_t* edge_pairs;          /* [E,2] (edge, shape) */
    const int32_t* edge_indices;        /* [edges,4] (o0,o1,v0,v1) */
    int32_t face_pair_count;
    const int32_t* face_pairs;          /* [F,2] (triangle, shape) */
    const int32_t* tri_indices;         /* [T,3] */
} VbdCollideDesc;

const char* vbd_last_error(void);
int32_t vbd_version(void);

/* SolverVBD.__init__ (particle part): stores the model pointers (the caller
 * keeps them alive), allocates every scratch buffer, reads the static model
 * arrays back once and builds the color-pass layout (element records packed
 * per workgroup batch).  Synchronises the device; not for use during capture. */
int32_t vbd_create(const VbdModelDesc* model, const VbdSolverParams* params, VbdContext** out);
int32_t vbd_destroy(VbdContext* ctx);

/* Rebuilds the color-pass layout after the caller changed static model arrays
 * (element materials, rest data, flags, masses).  SolverVBD.notify_model_changed
 * (newton/_src/solvers/solver.py:375-410, solver_vbd.py:851-856).  When the new layout has the
 * shape of the old one (same array sizes, per-color batch ranges and batch width: the case of
 * material and rest-data edits that leave every element's stiffness non-zero) it is written into
 * the existing device buffers, so graphs captured before stay valid and see the new values.
 * Otherwise (elements switched to or from zero stiffness, particles activated or deactivated,
 * new color groups) the layout is re-allocated and graphs captured before this call must be
 * captured again.  Synchronises the device; not for use during capture. */
int32_t vbd_rebuild(VbdContext* ctx);

/* Scalar settings read per launch: iterations of vbd_step, friction_epsilon, and the model's
 * soft_contact_ke / kd / mu (averaged with each shape's material in vbd_contact_prepare).  The
 * reference reads SolverVBD.iterations / friction_epsilon and Model.soft_contact_* at every step
 * (solver_vbd.py:1773-1825, :2493-2514); graphs captured before keep the values they were captured
 * with. */
int32_t vbd_set_scalars(VbdContext* ctx, int32_t iterations, float friction_epsilon, float soft_contact_ke,
                        float soft_contact_kd, float soft_contact_mu);

/* Layout statistics: out4 = {fast batches, generic-path vertices, packed element items, colors}. */
int32_t vbd_layout_stats(VbdContext* ctx, int64_t* out4);

/* Autotuner of the color-pass shape: builds the layout for 256-lane work-groups
 * (batches of <= 21 vertices / 256 element items), 192-lane ones (<= 16 / 192) and
 * 128-lane ones (<= 10 / 128), with item ranges packed or starting on wave
 * boundaries, times three force-assembly passes over every color with each
 * candidate -- 256 lanes at the compiler's register budget, 256 lanes at a
 * seven-waves-per-SIMD budget, 128 lanes, 256 lanes wave-aligned at either budget,
 * 192 lanes wave-aligned at either budget (q is only read; call after
 * vbd_contact_prepare and vbd_forward) -- and keeps the fastest.  block_out[3]
 * receives the lanes (256, 192 or 128), the waves target (1: compiler's choice, 7)
 * and the alignment (0 packed, 1 wave-aligned); ms_out[7] the seven per-pass-set
 * times.  Results do not depend on
 * the shape.  No counterpart in the reference; synchronises and allocates (not for
 * capture). */
int32_t vbd_autotune(VbdContext* ctx, void* stream, const float* q, float dt, int32_t* block_out, float* ms_out);

/* Register-budget occupancy target of the 256-lane color pass: 1 (compiler's
 * choice) or 7 waves per SIMD; what vbd_autotune chooses between. */
int32_t vbd_set_pass_waves(VbdContext* ctx, int32_t waves);

/* Item ranges of a color-pass batch packed (0) or each starting on a wave boundary (1), so that a
 * wave evaluates one element kind only; rebuilds the layout (synchronises, allocates); what
 * vbd_autotune also chooses between. */
int32_t vbd_set_aligned_items(VbdContext* ctx, void* stream, int32_t aligned);

/* Selects the work-group shape (256, 192 or 128 lanes) directly, rebuilding the
 * layout; the contact ranges of the current substep are carried over.  Not for
 * capture. */
int32_t vbd_set_block_shape(VbdContext* ctx, void* stream, int32_t block);

/* All three shape choices at once (lanes 256 / 192 / 128, waves 1 / 7, aligned 0 / 1), with one
 * layout rebuild; what vbd_autotune picks between.  Synchronises; not for capture. */
int32_t vbd_set_pass_shape(VbdContext* ctx, void* stream, int32_t block, int32_t waves, int32_t aligned);
/* out3 = {lanes, waves, aligned} of the current pass shape. */
int32_t vbd_get_pass_shape(VbdContext* ctx, int32_t* out3);

/* Full substep: contact prepare, forward, iterations x colors, finalize.
 * q_in is overwritten with the iterate (as the reference does); q_out/qd_out
 * receive the result.  contacts may be NULL (no soft contacts). */
int32_t vbd_step(VbdContext* ctx, void* stream, float* q_in, const float* qd_in, const float* f_ext,
                 float* q_out, float* qd_out, const VbdContactsDesc* contacts, float dt);

/* The pieces of vbd_step, for callers that interleave their own work. */
int32_t vbd_contact_prepare(VbdContext* ctx, void* stream, const VbdContactsDesc* contacts);
int32_t vbd_forward(VbdContext* ctx, void* stream, float* q, const float* qd, const float* f_ext, float dt);
int32_t vbd_iteration(VbdContext* ctx, void* stream, float* q, float dt);
int32_t vbd_color_pass(VbdContext* ctx, void* stream, int32_t color, float* q, float dt);
int32_t vbd_finalize(VbdContext* ctx, void* stream, const float* q, float* q_out, float* qd_out, float dt);

/* SolverBase.integrate_particles (newton/_src/solvers/solver.py:20-56, :297-330): one semi-implicit
 * Euler step v' = v + (f * w + g * [w > 0]) * dt, |v'| limited to v_max, x' = x + v' * dt, for the n
 * particles of x/v/f ([n,3]) with inverse masses w and flags (bit 0 ACTIVE).  Inactive particles get
 * x' = x and leave v_new unwritten.  g is the gravity row ([gravity_count,3]) of the particle's world
 * (world may be NULL: world 0; -1 selects world 0).  Needs no context; enqueued on stream. */
int32_t vbd_integrate_particles(void* stream, int32_t n, const float* x, const float* v, const float* f,
                                const float* inv_mass, const int32_t* flags, const int32_t* world,
                                const float* gravity, int32_t gravity_count, float dt, float v_max, float* x_new,
                                float* v_new);

/* Incremental-potential force residual at q: g_i = m_i/h^2 (y_i - q_i) +
 * elastic + damping + contact forces (the force the block solve assembles),
 * zero for inactive particles.  g_out (optional) receives [N,3]; norm_sq_out
 * (device, 1 double) receives sum |g_i|^2.  Uses the inertia and contact
 * state of the current substep (call after vbd_forward / vbd_contact_prepare).
 * The norm is reduced in a fixed order (bit-reproducible).  No counterpart in
 * the reference, which has no residual; used by iterate-to-tolerance drivers. */
int32_t vbd_residual(VbdContext* ctx, void* stream, const float* q, float dt, float* g_out, double* norm_sq_out);

/* Iterate from a guess (device [N,3]): for solved particles (ACTIVE,
 * inv_mass != 0) d = guess - x0, else d = 0, and q_dst = x0 + d, the form
 * forward_step writes its inertial guess in (particle_vbd_kernels.py:1350-1375).
 * commit != 0 also stores d as the solver's displacement, making q_dst the
 * iterate; commit == 0 fills a trial buffer for vbd_residual.  Warm-start hook
 * with no counterpart in the reference; call after vbd_forward. */
int32_t vbd_set_iterate(VbdContext* ctx, void* stream, float* q_dst, const float* guess, int32_t commit);

/* vbd_set_iterate for the particles p_begin <= i < p_end only (the others keep their iterate and
 * displacement): one world of a batched model.  No counterpart in the reference. */
int32_t vbd_set_iterate_range(VbdContext* ctx, void* stream, float* q_dst, const float* guess, int32_t commit,
                              int32_t p_begin, int32_t p_end);

/* Runs VBD iterations on q until the squared residual norm is <= target_norm_sq
 * or max_iterations were run, evaluating the residual before the first
 * iteration and after every check_every iterations (host-synchronous: not for
 * graph capture).  Reports the iterations run and the last squared norm. */
int32_t vbd_iterate_to_residual(VbdContext* ctx, void* stream, float* q, float dt, double target_norm_sq,
                                int32_t max_iterations, int32_t check_every, int32_t* iterations_out,
                                double* norm_sq_out);

/* Per-world residuals for models holding several worlds in contiguous particle ranges:
 * norm_sq_out[w] (device, world_count doubles) = sum of |g_i|^2 over the particles
 * world_offsets[w] <= i < world_offsets[w+1] (device int32 [world_count+1]), each reduced
 * in a fixed order.  Batched iterate-to-tolerance drivers judge every world by its own
 * residual; no counterpart in the reference. */
int32_t vbd_residual_worlds(VbdContext* ctx, void* stream, const float* q, float dt, int32_t world_count,
                            const int32_t* world_offsets, double* norm_sq_out);

/* vbd_iterate_to_residual over worlds: each world is iterated until its squared norm is
 * <= target_norm_sq[w] (host array; evaluated every check_every iterations), then frozen --
 * its vertices are neither solved nor re-evaluated, so its iterate is the one a solve of
 * that world alone stops at -- until at most stop_at_live worlds are still live (0: all
 * frozen) or max_iterations were run.  The freezing is decided on the device (with
 * stop_at_live > 0 the remaining worlds are frozen as soon as at most that many are live); the
 * host polls the live-world count every 8 evaluations.  particle_world (device int32 [N]) gives each particle's world, consistent
 * with world_offsets; world_scratch_dev (2 * world_count doubles) and world_state_dev
 * (2 * world_count + 1 int32) are device scratch.  Reports the iterations run, optionally
 * each world's count (world_iterations_out, host), and the last norms (norm_sq_out, host).
 * Host-synchronous; not available with self-contact. */
int32_t vbd_iterate_to_residual_worlds(VbdContext* ctx, void* stream, float* q, float dt, int32_t world_count,
                                       const int32_t* world_offsets, const int32_t* particle_world,
                                       const double* target_norm_sq, int32_t max_iterations, int32_t check_every,
                                       int32_t stop_at_live, double* world_scratch_dev, int32_t* world_state_dev,
                                       int32_t* iterations_out, int32_t* world_iterations_out, double* norm_sq_out);

/* Copies the solver's scratch state of the current substep ([N,3] floats)
 * into dst (device) on the stream: which = 0 inertia y, 1 q_prev, 2 displacement. */
int32_t vbd_copy_scratch(VbdContext* ctx, void* stream, int32_t which, float* dst);

/* Restores the displacement ([N,3] floats, device) saved with vbd_copy_scratch(which = 2) together
 * with the positions it belongs to, so a saved iterate (q, d) continues bit for bit; rewriting q
 * through vbd_set_iterate would store d = fl(q - x0) instead, which may differ from d in its last
 * bits.  No counterpart in the reference (warm-start drivers). */
int32_t vbd_set_displacement(VbdContext* ctx, void* stream, const float* src);

/* Cloth self-contact (contexts created with params.self_contact = 1).  With it on, vbd_forward
 * detects before the forward step and truncates after it, vbd_iteration re-detects on the
 * collision-detection interval and truncates after every color pass (solver_vbd.py:2003-2063,
 * :2540-2545, :3221-3237), and the color pass adds each vertex's self-contact terms after its
 * springs (particle_vbd_kernels.py:1498-1641).
 *
 * TriMeshCollisionDetector.refit + vertex_triangle_collision_detection + edge_edge_collision_detection
 * (tri_mesh_collision.py:604-734) at positions q with the solver's margin; also makes q the base of
 * the displacements. */
int32_t vbd_self_detect(VbdContext* ctx, void* stream, const float* q);
/* Copies the last detection's results into caller buffers (each may be null): per vertex the
 * colliding triangles [N, vertex buffer size] (-1 padded, ascending), their count and the minimum
 * distance; per triangle the minimum distance to any colliding vertex; per edge the colliding
 * edges [E, edge buffer size], count and minimum distance. */
int32_t vbd_self_copy_lists(VbdContext* ctx, void* stream, int32_t* vt_list, int32_t* vt_count, float* vt_min,
                            float* tri_min, int32_t* ee_list, int32_t* ee_count, float* ee_min);
/* Synchronises the device; *grid_overflow = 1 when a primitive spanned more detection-grid cells
 * than a detection can list (its contacts were not detected). */
int32_t vbd_self_status(VbdContext* ctx, int32_t* grid_overflow);

/* Soft-contact generation (CollisionPipeline.collide, soft part): particle records in pair order, then
 * the full-surface edge and face records (soft_contacts_sdf.py:380-469, :277-377) in pair order;
 * soft_contact_count counts every hit (it may exceed soft_contact_max, as the reference's counter does).
 * vbd_collide_create reads desc.shape_body back once.  vbd_collide_generate is for static shapes only
 * (VBD_ERR_ARG when a shape has a body); vbd_collide_generate_bodies places body-attached shapes with
 * body_q ([body_count,7], the collided state's State.body_q, geometry/kernels.py:1065-1081). */
int32_t vbd_collide_create(const VbdCollideDesc* desc, VbdCollide** out);
int32_t vbd_collide_destroy(VbdCollide* col);
int32_t vbd_collide_generate(VbdCollide* col, void* stream, const float* particle_q, float margin,
                             const VbdContactsOut* out);
int32_t vbd_collide_generate_bodies(VbdCollide* col, void* stream, const float* particle_q, const float* body_q,
                                    int32_t body_count, float margin, const VbdContactsOut* out);

/* Largest pair_count vbd_collide_generate_noalloc accepts. */
#define VBD_NOALLOC_MAX_PAIRS (16 * 1024 * 1024)

/* vbd_collide_generate that allocates nothing: no counterpart in the reference.  It is for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation), and needs no VbdCollide: the descriptor is read per call.  Same records, same
 * order and same soft_contact_count as vbd_collide_generate.  Its device scratch (one int per
 * 1,024 pairs, 64 KiB) is a static array each GPU holds until the process ends; calls that overlap
 * on one GPU share it and race.  More than VBD_NOALLOC_MAX_PAIRS candidate pairs (particle, edge and
 * face pairs together) are rejected with VBD_ERR_ARG, and so is a non-NULL desc.shape_body (shapes
 * are static here).  The first call on a device looks the scratch address up (not inside a capture). */
int32_t vbd_collide_generate_noalloc(const VbdCollideDesc* desc, void* stream, const float* particle_q, float margin,
                                     const VbdContactsOut* out);

#ifdef __cplusplus
}
#endif

#endif
