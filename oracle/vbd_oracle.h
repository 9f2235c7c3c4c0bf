/*
 * CPU oracle for the particle Vertex Block Descent step -- TEST INFRASTRUCTURE.
 *
 * This library is the checker, never the product: only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it.  It restates the
 * reference algorithm (Newton's SolverVBD particle path, Warp CPU ordering)
 * in plain C, instantiated twice:
 *   *_f32  -- FP32 "reference semantics" oracle (same op order as the
 *             reference's CPU kernels, no FMA contraction),
 *   *_f64  -- FP64 accuracy oracle of the same iteration.
 *
 * Model arrays are the reference's FP32/INT32 layouts (AoS vec3, row-major
 * mat22/mat33).  State and contact arrays use the instantiation's REAL type.
 */
#ifndef VBD_ORACLE_H
#define VBD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleModel {
    int32_t particle_count;
    int32_t tri_count;
    int32_t edge_count;
    int32_t tet_count;
    const float* particle_mass;      /* [N] */
    const float* particle_inv_mass;  /* [N] */
    const float* particle_radius;    /* [N] */
    const int32_t* particle_flags;   /* [N] bit0 = ACTIVE */
    const int32_t* particle_colors;  /* [N] */
    const int32_t* particle_world;   /* [N] */
    const int32_t* tri_indices;      /* [T*3] */
    const float* tri_poses;          /* [T*4] row-major inverse rest basis */
    const float* tri_materials;      /* [T*5] (ke=mu, ka=lambda, kd, drag, lift) */
    const float* tri_areas;          /* [T] */
    const int32_t* edge_indices;     /* [E*4] (o0, o1, v0, v1) */
    const float* edge_rest_angle;    /* [E] */
    const float* edge_rest_length;   /* [E] */
    const float* edge_bending_properties; /* [E*2] (ke, kd) */
    const int32_t* tet_indices;      /* [K*4] */
    const float* tet_poses;          /* [K*9] row-major Dm^-1 */
    const float* tet_materials;      /* [K*3] (mu, lambda, kd) */
    const int32_t* v_adj_tris;       /* CSR (element, slot) pairs */
    const int32_t* v_adj_tris_offsets;
    const int32_t* v_adj_edges;
    const int32_t* v_adj_edges_offsets;
    const int32_t* v_adj_tets;
    const int32_t* v_adj_tets_offsets;
    int32_t color_count;
    const int32_t* color_offsets;    /* [C+1] into color_particles */
    const int32_t* color_particles;  /* concatenated color groups */
    float gravity[3];
    int32_t shape_count;
    const int32_t* shape_type;       /* GeoType: 1 plane, 3 sphere */
    const int32_t* shape_body;       /* -1 = static */
    const int32_t* shape_flags;      /* bit2 = COLLIDE_PARTICLES */
    const int32_t* shape_world;
    const float* shape_transform;    /* [S*7] (px,py,pz, qx,qy,qz,qw) */
    const float* shape_scale;        /* [S*3] */
    const float* shape_margin;       /* [S] */
    const float* shape_material_ke;
    const float* shape_material_kd;
    const float* shape_material_mu;
    int32_t world_count;
    float soft_contact_ke;
    float soft_contact_kd;
    float soft_contact_mu;
    int32_t spring_count;
    const int32_t* spring_indices;   /* [S*2] */
    const float* spring_rest_length; /* [S] */
    const float* spring_stiffness;   /* [S] */
    const float* spring_damping;     /* [S] */
} OracleModel;

struct OracleSelf;

typedef struct OracleParams {
    int32_t iterations;
    float friction_epsilon;
    float avbd_beta;          /* rigid_avbd_linear_beta; 0 = fixed penalty */
    float k_start;            /* < 0 => penalty seeded at the averaged ke */
    int32_t use_libm_atan2;   /* 1: C library atan2f, 0: portable atan2 (matches the HIP kernels) */
    int32_t fp64_ill_conditioned; /* 1: solve blocks with cond > 1e6 in double */
    int32_t num_threads;      /* OpenMP threads for the per-color vertex loop (<=1: serial) */
    const struct OracleSelf* self_contact; /* NULL: particle_enable_self_contact = False */
    double residual_stop;     /* > 0 with a residual history: stop once residual <= residual_stop * initial residual */
    /* poses of the bodies that shapes are attached to (one-way coupling to an external rigid solver,
     * solver_vbd.py:2523-2526): body_q = this step's (state_out.body_q), body_q_prev = the step's start
     * (state_in.body_q); [body_count * 7] (p, q xyzw); NULL when every shape is static */
    const float* body_q;
    const float* body_q_prev;
    int32_t body_count;
} OracleParams;

/* Cloth self-contact settings (SolverVBD particle_* arguments, solver_vbd.py:231-242) and the
 * optional outputs of the last collision detection.  Triangle / edge lists are stored in
 * ascending primitive order (the reference records them in BVH traversal order; the two agree
 * whenever a buffer does not overflow). */
typedef struct OracleSelf {
    float radius;                    /* particle_self_contact_radius */
    float margin;                    /* particle_self_contact_margin (detection query radius) */
    float relaxation;                /* particle_conservative_bound_relaxation (gamma) */
    float max_displacement;          /* float32(margin * relaxation * 0.5), computed in double */
    int32_t vt_capacity;             /* particle_vertex_contact_buffer_size */
    int32_t ee_capacity;             /* particle_edge_contact_buffer_size */
    int32_t detection_interval;      /* particle_collision_detection_interval */
    float edge_parallel_eps;         /* particle_edge_parallel_epsilon */
    float rest_exclusion_radius;     /* particle_rest_shape_contact_exclusion_radius */
    const float* q_rest;             /* [N*3] model.particle_q (rest-shape exclusion) */
    const int32_t* vt_filter;        /* per-vertex sorted triangle ids to skip (NULL: none) */
    const int32_t* vt_filter_offsets;/* [N+1] */
    const int32_t* ee_filter;        /* per-edge sorted edge ids to skip (NULL: none) */
    const int32_t* ee_filter_offsets;/* [E+1] */
    /* outputs of the last detection (each may be NULL) */
    int32_t* vt_list;                /* [N*vt_capacity], -1 padded */
    int32_t* vt_count;               /* [N] (may exceed the capacity) */
    double* vt_min_dist;             /* [N] */
    double* tri_min_dist;            /* [T] */
    int32_t* ee_list;                /* [E*ee_capacity], -1 padded */
    int32_t* ee_count;               /* [E] */
    double* ee_min_dist;             /* [E] */
} OracleSelf;

/* Soft-contact candidate pairs: [P*2] (particle, shape) for create_soft_contacts, and for the
 * full-surface passes [E*2] (edge, shape) and [F*2] (triangle, shape) (counts 0 when off). */
typedef struct OraclePairs {
    int32_t pair_count;
    const int32_t* pairs;
    int32_t edge_pair_count;
    const int32_t* edge_pairs;
    int32_t face_pair_count;
    const int32_t* face_pairs;
} OraclePairs;

#define ORACLE_DECLARE(SFX, REAL)                                                              \
    int32_t vbd_oracle_step_##SFX(const OracleModel* m, const OracleParams* p,                 \
                                  int32_t contact_count, int32_t contact_max,                  \
                                  const int32_t* c_indices, const int32_t* c_shape,            \
                                  const REAL* c_body_pos, const REAL* c_body_vel,              \
                                  const REAL* c_normal, const REAL* c_bary, REAL* q_in,        \
                                  const REAL* qd_in, const REAL* f_ext, REAL* q_out,           \
                                  REAL* qd_out, REAL dt, double* residual_hist);               \
    int32_t vbd_oracle_create_soft_contacts_##SFX(                                             \
        const OracleModel* m, const OraclePairs* pairs, const REAL* q, const float* body_q,    \
        int32_t body_count, REAL margin, int32_t contact_max, int32_t* c_count,                \
        int32_t* c_particle, int32_t* c_indices, REAL* c_bary, int32_t* c_shape,               \
        REAL* c_body_pos, REAL* c_body_vel, REAL* c_normal);                                   \
    int32_t vbd_oracle_simulate_##SFX(const OracleModel* m, const OracleParams* p,             \
                                      const OraclePairs* pairs, REAL* q, REAL* qd,             \
                                      const REAL* f_ext, int32_t substeps, REAL dt,            \
                                      REAL margin, int32_t contact_max, int32_t record_every,  \
                                      REAL* q_record);                                         \
    void vbd_oracle_membrane_##SFX(const REAL* x /*9*/, const REAL* x_anchor /*9*/,           \
                                   const float* pose /*4*/, float area, float mu, float lmbd,  \
                                   float kd, REAL dt, int32_t v_order, REAL* f, REAL* h);      \
    void vbd_oracle_bending_##SFX(const REAL* x /*12*/, const REAL* x_anchor /*12*/,          \
                                  float rest_angle, float rest_length, float ke, float kd,     \
                                  REAL dt, int32_t v_order, int32_t use_libm_atan2, REAL* f,   \
                                  REAL* h);                                                    \
    void vbd_oracle_tet_##SFX(const REAL* x /*12*/, const REAL* x_prev /*12*/,                \
                              const float* dm_inv /*9*/, float mu, float lmbd, float kd,       \
                              REAL dt, int32_t v_order, REAL* f, REAL* h);                     \
    void vbd_oracle_body_contact_##SFX(const REAL* x, const REAL* x_prev, const REAL* body_pos,\
                                       const REAL* body_vel, const REAL* normal, REAL radius,  \
                                       REAL margin, REAL ke, REAL kd, REAL mu,                 \
                                       REAL friction_epsilon, REAL dt, REAL* f, REAL* h);      \
    int32_t vbd_oracle_block_solve_##SFX(const REAL* h /*9*/, const REAL* f /*3*/,            \
                                         int32_t fp64_ill_conditioned, REAL* dx /*3*/);        \
    REAL vbd_oracle_atan2_##SFX(REAL y, REAL x);                                               \
    void vbd_oracle_spring_##SFX(const REAL* x /*6: x0, x1*/, const REAL* anchor /*6*/,        \
                                 REAL rest, REAL ke, REAL kd, REAL dt, REAL* f /*6: f0, f1*/,    \
                                 REAL* h /*9*/);                                                  \
    int32_t vbd_oracle_triangle_closest_##SFX(const REAL* abcp /*12*/, REAL* closest /*3*/,   \
                                              REAL* bary /*3*/);                              \
    void vbd_oracle_closest_edge_edge_##SFX(const REAL* x /*12*/, REAL eps, REAL* std /*3*/);  \
    void vbd_oracle_self_force_norm_##SFX(REAL dis, REAL radius, REAL k, REAL* out /*2*/);      \
    int32_t vbd_oracle_vt_contact_##SFX(const REAL* abcp /*12*/, const REAL* anchor /*12*/,     \
                                        REAL radius, REAL ke, REAL kd, REAL mu,                \
                                        REAL friction_epsilon, REAL dt, REAL* f /*12*/,         \
                                        REAL* h /*36*/);                                       \
    int32_t vbd_oracle_ee_contact_##SFX(const REAL* x /*12*/, const REAL* anchor /*12*/,        \
                                        REAL radius, REAL ke, REAL kd, REAL mu,                \
                                        REAL friction_epsilon, REAL dt, REAL parallel_eps,     \
                                        REAL* f /*6*/, REAL* h /*18*/);                         \
    void vbd_oracle_truncation_ts_##SFX(int32_t kind, const REAL* x /*12*/,                   \
                                        const REAL* delta /*12*/, REAL parallel_eps,           \
                                        REAL gamma, REAL* t /*4*/);                              \
    int32_t vbd_oracle_self_detect_##SFX(const OracleModel* m, const struct OracleSelf* sc,    \
                                         const REAL* q);                                       \
    int32_t vbd_oracle_residual_##SFX(const OracleModel* m, const OracleParams* p,             \
                                      int32_t contact_count, int32_t contact_max,              \
                                      const int32_t* c_indices, const int32_t* c_shape,        \
                                      const REAL* c_body_pos, const REAL* c_body_vel,          \
                                      const REAL* c_normal, const REAL* c_bary, const REAL* q, \
                                      const REAL* q_prev, const REAL* inertia, REAL dt,        \
                                      REAL* g_out, double* norm_out);

ORACLE_DECLARE(f32, float)
ORACLE_DECLARE(f64, double)

#ifdef __cplusplus
}
#endif

#endif
