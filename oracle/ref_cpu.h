/*
 * ref_cpu — CPU restatement of the dyphur frame step, used ONLY as test
 * infrastructure: the parity oracle for the gfx950 kernels and the host-CPU
 * baseline timed beside them.  Nothing in the product links, loads or calls
 * this library; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do.
 *
 * It restates the reference algorithms stage by stage (same loop structure,
 * same floating-point operation order, one OpenMP parallel-for per reference
 * parallel_for, including the multi-launch bitonic sorts), compiled with
 * -O2 -ffp-contract=off and no fast-math.  Reference files followed:
 *   core/src/integrator.cpp:7-71          (rc_integrate)
 *   core/src/broadphase.cpp:14-365        (rc_bp_*)
 *   core/src/narrowphase.cpp:12-306,912-1127 (rc_np_run)
 *   core/src/xpbd_solver.cpp:1-134,409-521,558-733 (rc_solve)
 *   core/src/xpbd_solver.cpp:60-407,523-556 (joints: rc_solve_joints)
 *   core/src/body_store.cpp:28-69, core/include/core/math/mat3.hpp:61-79
 *                                          (rc_body_setup)
 *   compute/include/compute/sort.hpp:68-85 (rc_bitonic_sort_*)
 *   examples/stress_test_blocks/main.cpp:34-41,301-311 (rc_fnv1a64_state)
 *   core/src/ray_query.cpp:12-195          (rc_ray_query)
 *
 * Semantic choices at the SYCL boundary (AdaptiveCpp is not available here):
 *   sycl::half(float) = IEEE binary16 round-to-nearest-even;
 *   sycl::rsqrt(x)    = 1.f / sqrtf(x) (correctly rounded sqrt and divide);
 *   sycl::clz(0ull)   = 64;
 *   sycl::atan2(y, x) = the correctly rounded float atan2 (the C library's
 *                       float atan2 under rc_set_libc_atan2(1)).
 *
 * Pinning (tests/test_oracle_kats.py, tests/test_oracle_pins.py): the
 * reference's KATs and golden hash, brute force (Morton keys and order,
 * Karras tree, pair set), closed forms (leaf boxes, integrator), a direct
 * simulation of the JP round rule, invariants.  Parity unpinned: the
 * box-box manifold heuristic, reference-face swap, clip-vertex order and
 * centroid bits for general poses (core/src/narrowphase.cpp:216-305), and
 * the exact bits of multi-contact solves beyond the reference's solver KATs
 * (core/src/xpbd_solver.cpp:413-521,559-587); only a run of the reference
 * could pin those, and it cannot be built without SYCL.
 */
#ifndef DYPHUR_REF_CPU_H
#define DYPHUR_REF_CPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host SoA body arrays; same field order as the reference BodyView
 * (core/include/core/body.hpp:30-46). */
typedef struct rc_bodies {
    float *pos_x, *pos_y, *pos_z;
    float *rot_w, *rot_x, *rot_y, *rot_z;
    float *vel_x, *vel_y, *vel_z;
    float *ang_x, *ang_y, *ang_z;
    float *inv_mass;
    float *iI_xx, *iI_yy, *iI_zz;
    float *iI_xy, *iI_xz, *iI_yz;
    uint32_t *shape, *flags;
    uint32_t n;
} rc_bodies;

/* Shape catalog (core/include/core/shapes.hpp:22-28). */
typedef struct rc_shapes {
    const uint32_t *type;
    const float *half_x, *half_y, *half_z;
    const uint32_t *ext_id;
    uint32_t n;
} rc_shapes;

typedef struct rc_aabb { float min_x, min_y, min_z, max_x, max_y, max_z; } rc_aabb;
typedef struct rc_pair { uint32_t a, b; } rc_pair;

/* Host views of the oracle's contact store (core/include/core/contact.hpp). */
typedef struct rc_contacts {
    const uint32_t *body_a, *body_b;
    const float *pos_x, *pos_y, *pos_z;
    const float *norm_x, *norm_y, *norm_z;
    const float *depth;
    uint32_t n;          /* stored (clamped) count */
    uint32_t last_count; /* true count before clamping */
} rc_contacts;

typedef struct rc_solve_stats {
    uint32_t n_constraints;
    uint32_t used_serial;  /* 1 = run_serial path taken */
    uint32_t n_colors;     /* colours used by the coloured path */
    uint32_t rounds;       /* JP rounds executed */
    uint32_t overflow;     /* >64 colours on a body */
} rc_solve_stats;

/* ---- numeric building blocks ------------------------------------------ */
uint16_t rc_half_from_float(float f);
float    rc_half_to_float(uint16_t h);
void     rc_half_from_float_array(const float *in, uint16_t *out, size_t n);
uint32_t rc_hash_u32(uint32_t x);
uint32_t rc_morton3(uint32_t x, uint32_t y, uint32_t z);
uint64_t rc_fnv1a64(uint64_t h, const void *data, size_t n);
uint64_t rc_fnv1a64_state(const rc_bodies *b);

/* BodyStore::add for n bodies: mass[n], inertia[9n] row-major, flags_in[n]
 * -> inv_mass, iI_* (inverse inertia, zero for statics), flags (Static OR-ed). */
void rc_body_setup(uint32_t n, const float *mass, const float *inertia,
                   const uint32_t *flags_in, rc_bodies *out);

/* ---- stages -------------------------------------------------------------*/
void rc_integrate(rc_bodies *b, float gx, float gy, float gz, float dt, int n_substeps);

void rc_bitonic_sort_u64_u32(uint64_t *keys, uint32_t *vals, size_t n);
void rc_bitonic_sort_u32_u32(uint32_t *keys, uint32_t *vals, size_t n);
void rc_bitonic_sort_u32(uint32_t *keys, size_t n);

typedef struct rc_world rc_world;
rc_world *rc_world_create(uint32_t max_bodies, uint32_t max_pairs,
                          uint32_t max_contacts, int n_iters, float mu);
void rc_world_destroy(rc_world *w);

/* Broadphase::build_and_query + download_count. Returns the true pair count. */
uint32_t rc_bp_build_and_query(rc_world *w, const rc_bodies *b, const rc_shapes *s,
                               const rc_aabb *bounds);
/* Copies BVH internals out (any pointer may be NULL).  aabb holds 8 u16 per
 * node (mn xyz, mx xyz, pad, pad); node count 2n-1; internal n-1. */
void rc_bp_copy_out(const rc_world *w, uint32_t n, uint64_t *morton, uint32_t *sorted_idx,
                    int32_t *left, int32_t *right, int32_t *parent, int32_t *root,
                    uint16_t *aabb, uint32_t *refit_rounds);
void rc_bp_sort_pairs(rc_world *w, uint32_t n_pairs);
void rc_bp_copy_pairs(const rc_world *w, rc_pair *out, uint32_t n);
const rc_pair *rc_bp_pairs(const rc_world *w);

/* Narrowphase::run (CPU pair count). Returns stored contact count. */
uint32_t rc_np_run(rc_world *w, const rc_pair *pairs, uint32_t n_pairs,
                   const rc_bodies *b, const rc_shapes *s);
void rc_np_contacts(const rc_world *w, rc_contacts *out);
/* Feature words of the stored contacts (first n, in store order).  The word
 * (no reference counterpart; narrowphase.cpp:153-305 computes its parts and
 * keeps only the contact):
 *   bits 0-2   kind: 0 box-box, face of A is the reference face; 1 box-box,
 *              face of B; 2 box-box edge-edge; 3 sphere-sphere; 4 sphere-box,
 *              centre outside (closest point); 5 sphere-box, centre inside
 *   bits 3-4   reference-face axis (0, 1), A's edge axis (2), the face of least
 *              penetration (5)
 *   bits 5-6   B's edge axis (2)
 *   bits 8-15  kinds 0/1: incident vertices k = 0..7 (signs of the half
 *              extents by the bits of k) that passed the depth and both side
 *              tests; the first four in k order form the manifold
 *   bits 16-19 box-box: the SAT axis of least overlap before the reference-face
 *              choice (0-2 A's faces, 3-5 B's faces, 6 + 3i + j edge i x j) */
void rc_np_features(const rc_world *w, uint32_t *out, uint32_t n);
/* Overwrite the contact store (for solver-only tests). */
void rc_np_set_contacts(rc_world *w, uint32_t n, const uint32_t *body_a, const uint32_t *body_b,
                        const float *px, const float *py, const float *pz,
                        const float *nx, const float *ny, const float *nz, const float *depth);

/* Joint SoA, field order of the reference JointView
 * (core/include/core/articulation.hpp:47-70); type 0 fixed, 1 revolute,
 * 2 prismatic, 3 ball. */
typedef struct rc_joints {
    const uint32_t *body_parent, *body_child;
    const float *anchor_px, *anchor_py, *anchor_pz;
    const float *anchor_cx, *anchor_cy, *anchor_cz;
    const float *axis_px, *axis_py, *axis_pz;
    const float *limit_lo, *limit_hi;
    const float *target_pos, *target_vel;
    const float *stiffness, *damping;
    const float *compliance_pos, *compliance_ang;
    const uint8_t *type;
    uint32_t n;
} rc_joints;

/* XpbdSolver::solve with JointView{} (n = 0). */
void rc_solve(rc_world *w, rc_bodies *b, float dt, int force_serial, rc_solve_stats *stats);
/* XpbdSolver::solve(s, contacts, joints, bodies, dt) (xpbd_solver.cpp:590-733):
 * constraint k is contact k (k < nc) or joint k - nc; joints may be NULL. */
void rc_solve_joints(rc_world *w, rc_bodies *b, const rc_joints *joints, float dt, int force_serial,
                     rc_solve_stats *stats);
/* The atan2 of the revolute joint angle: correctly rounded (default), or the
 * C library's float atan2 (on = 1), which the reference itself calls. */
void rc_set_libc_atan2(int on);
float rc_joint_atan2(float y, float x);
/* Colour array of the last coloured solve (n_constraints entries). */
void rc_solve_colors(const rc_world *w, uint32_t *out, uint32_t n);

/* One frame of the stress driver loop (examples/stress_test_blocks/main.cpp:265-272):
 * integrate -> build_and_query -> download_count -> sort_pairs -> narrowphase -> solve. */
typedef struct rc_frame_stats { uint32_t n_pairs, n_contacts, n_colors, used_serial; } rc_frame_stats;
void rc_frame(rc_world *w, rc_bodies *b, const rc_shapes *s, const rc_aabb *bounds,
              float gx, float gy, float gz, float dt, int n_substeps, int force_serial,
              rc_frame_stats *st);
/* rc_frame with joints (NULL = none). */
void rc_frame_joints(rc_world *w, rc_bodies *b, const rc_shapes *s, const rc_aabb *bounds, const rc_joints *joints,
                     float gx, float gy, float gz, float dt, int n_substeps, int force_serial,
                     rc_frame_stats *stats);

/* ---- per-element entry points (for function-level parity checks) ----- */
/* Contact of one pair (narrowphase.cpp:920-1000): ha/hb half extents, pa/pb
 * centres, qa/qb quaternions wxyz.  out[7] = p xyz, n xyz, depth; returns 1 on hit. */
int rc_pair_contact_feat(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa,
                         const float *pb, const float *qa, const float *qb, float *out, uint32_t *feat);
int rc_pair_contact(uint32_t ta, uint32_t tb, const float *ha, const float *hb, const float *pa,
                    const float *pb, const float *qa, const float *qb, float *out);
/* One solve_contact (xpbd_solver.cpp:413-521) on two bodies given as 20 floats
 * (pos xyz, rot wxyz, vel xyz, ang xyz, inv_mass, iI xx yy zz xy xz yz). */
void rc_project_contact(const float *contact, float *a, float *b, float mu, float *lam, float dt);
/* Leaf AABB half extents of one body (broadphase.cpp:222-244). */
void rc_leaf_extent(uint32_t stype, const float *half3, const float *q4, float *out3);
/* Karras node i (broadphase.cpp:164-174): out = lo, hi, split, left, right. */
void rc_karras_node(const uint64_t *codes, int n, int i, int32_t *out);

/* Ray queries over the last build_and_query's LBVH (core/src/ray_query.cpp:126-195,
 * core/include/core/ray_query.hpp:9-18): closest exact hit (oriented box slab test,
 * analytic sphere) with t < t_max; body 0xFFFFFFFF and t = t_max when none. */
typedef struct rc_ray { float ox, oy, oz, dx, dy, dz, t_max; } rc_ray;
typedef struct rc_ray_hit { float t; uint32_t body; } rc_ray_hit;
void rc_ray_query(const rc_world *w, uint32_t n, const rc_bodies *b, const rc_shapes *s, const rc_ray *rays,
                  rc_ray_hit *hits, uint32_t n_rays);

int rc_max_threads(void);
void rc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
