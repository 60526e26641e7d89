/*
 * dyphur_hip.h — C ABI of the gfx950 frame step (libdyphur_hip.so).
 *
 * The reference exposes its path as C++ classes over a SYCL queue; this ABI
 * sits beneath an unchanged C++ surface (compute::Device/Stream/Buffer,
 * core::Broadphase/Narrowphase/XpbdSolver/integrate) and beneath the Python
 * module.  Plain pointers and sizes only.  Every function returns an int
 * status: 0 on success, a positive hipError_t from the runtime, or one of the
 * negative DY_ERR_* codes below.
 *
 * Memory: all dy_*_view pointers are device pointers obtained from dy_malloc
 * (or any hipMalloc'd memory).  Stage objects own their scratch and are bound
 * to one stream; they are not re-entrant.
 *
 * Blocking calls (host waits for the stream): dy_memcpy_d2h, dy_stream_synchronize,
 * dy_event_synchronize, dy_reduce_sum_f32, *_download_count, dy_narrowphase_last_count,
 * dy_broadphase_get_timing,
 * dy_state_hash_fnv1a64, dy_xpbd_last_stats, dy_xpbd_get_timing,
 * dy_xpbd_download_classes, dy_batch_stats, dy_batch_autotune.  Everything else,
 * dy_xpbd_solve included, is asynchronous on the given stream, in stream order.
 */
#ifndef DYPHUR_HIP_H
#define DYPHUR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DY_OK 0
#define DY_ERR_INVALID_ARG (-1)  /* null/oversized argument, capacity exceeded */
#define DY_ERR_UNSUPPORTED (-2)  /* convex-hull or triangle-mesh shapes */
#define DY_ERR_NO_DEVICE   (-3)  /* no HIP device */
#define DY_ERR_INTERNAL    (-4)  /* a device-side consistency check failed */

/* Body flags: core/include/core/body.hpp:8-12 */
#define DY_BODY_STATIC    1u
#define DY_BODY_KINEMATIC 2u
#define DY_BODY_SLEEPING  4u

/* Shape types: core/include/core/shapes.hpp:6-11 (2 and 3 are rejected) */
#define DY_SHAPE_BOX    0u
#define DY_SHAPE_SPHERE 1u

/* SoA body view; layout identical to dyphur::BodyView (core/include/core/body.hpp:30-46). */
typedef struct dy_body_view {
    float *pos_x, *pos_y, *pos_z;
    float *rot_w, *rot_x, *rot_y, *rot_z;
    float *vel_x, *vel_y, *vel_z;
    float *ang_x, *ang_y, *ang_z;
    float *inv_mass;
    float *iI_xx, *iI_yy, *iI_zz;
    float *iI_xy, *iI_xz, *iI_yz;
    uint32_t *shape;
    uint32_t *flags;
    uint32_t n;
} dy_body_view;

/* Shape catalog; layout identical to dyphur::ShapeView (core/include/core/shapes.hpp:22-28). */
typedef struct dy_shape_view {
    const uint32_t *type;
    const float *half_x, *half_y, *half_z;
    const uint32_t *ext_id;
    uint32_t n;
} dy_shape_view;

/* Contact store view; layout identical to dyphur::ContactView (core/include/core/contact.hpp:8-20). */
typedef struct dy_contact_view {
    uint32_t *body_a, *body_b;
    float *pos_x, *pos_y, *pos_z;
    float *norm_x, *norm_y, *norm_z;
    float *depth;
    uint32_t *n;        /* device pointer: live contact count */
    uint32_t capacity;
} dy_contact_view;

/* JointView (core/include/core/articulation.hpp:47-70): device SoA arrays,
 * n joints.  type: 0 fixed, 1 revolute, 2 prismatic, 3 ball; anchors in the
 * parent's / child's body frame, axis in the parent's frame; limit_lo >
 * limit_hi disables the limits. */
typedef struct dy_joint_view {
    uint32_t *body_parent, *body_child;
    float *anchor_px, *anchor_py, *anchor_pz;
    float *anchor_cx, *anchor_cy, *anchor_cz;
    float *axis_px, *axis_py, *axis_pz;
    float *limit_lo, *limit_hi;
    float *target_pos, *target_vel;
    float *stiffness, *damping;
    float *compliance_pos, *compliance_ang;
    uint8_t *type;
    uint32_t n;
} dy_joint_view;

/* core/include/core/aabb.hpp:7-31 */
typedef struct dy_aabb { float min_x, min_y, min_z, max_x, max_y, max_z; } dy_aabb;

/* core/include/core/broadphase.hpp:12-14: a < b always. */
typedef struct dy_contact_pair { uint32_t a, b; } dy_contact_pair;

/* core/include/core/broadphase.hpp:21-25: 6 binary16 bounds + 2 zero pads, 16 B. */
typedef struct dy_node_box { uint16_t mn_x, mn_y, mn_z, mx_x, mx_y, mx_z, pad0, pad1; } dy_node_box;

/* core/include/core/broadphase.hpp:30-38: BVH internals for ray queries. */
typedef struct dy_bvh_view {
    const int32_t *left;        /* internal nodes [0, n-2] */
    const int32_t *right;       /* internal nodes [0, n-2] */
    const int32_t *parent;      /* all nodes [0, 2n-2], -1 for the root */
    const int32_t *root;        /* 1 element */
    const uint32_t *sorted_idx; /* leaf k -> body index */
    const dy_node_box *aabb;    /* all nodes [0, 2n-2] */
    uint32_t n;
} dy_bvh_view;

typedef struct dy_device_info {
    char name[256];
    int is_gpu;
    int arch_major, arch_minor; /* gcnArch, e.g. 9, 5 for gfx950 */
    size_t compute_units;
    size_t global_mem_bytes;
} dy_device_info;

typedef struct dy_solve_stats {
    uint32_t n_constraints;  /* constraints solved: contacts + joints */
    uint32_t used_serial;    /* 1: reference-order (run_serial) semantics */
    uint32_t n_classes;      /* colours (coloured path) or dependency levels (serial path) */
    uint32_t rounds;         /* colouring rounds the reference would run (max dependency depth) */
    uint32_t overflow;       /* 1: a body would need more than 64 colours -> serial fallback */
    uint32_t dyn_endpoints;  /* contact endpoints on dynamic bodies (sum of per-body contact counts) */
} dy_solve_stats;

/* Kernel timing of the solver, accumulated over solves while enabled: HIP
 * events recorded on the solve's stream around the class-assignment kernel and
 * around the sweep kernel. */
typedef struct dy_solver_timing {
    uint32_t solves;          /* solves timed */
    double assign_ms;         /* total class-assignment kernel time */
    double sweep_ms;          /* total sweep kernel time */
    double sweep_bytes;       /* total algorithmic bytes of the timed sweeps */
} dy_solver_timing;

/* Kernel timing of the broadphase, accumulated over builds while enabled: HIP
 * events recorded on the build's stream between its launches.  No reference
 * counterpart (the reference's profile mode fences whole stages:
 * examples/stress_test_blocks/main.cpp:198-257). */
typedef struct dy_broadphase_timing {
    uint32_t builds;          /* builds timed */
    double sort_ms;           /* Morton keys + key sort */
    double karras_ms;         /* Karras tree */
    double refit_ms;          /* fp16 leaf boxes + bottom-up refit */
    double traverse_ms;       /* traversal: pair counts and sorted rows */
    double emit_ms;           /* row-count scan + pair emission */
} dy_broadphase_timing;

typedef struct dy_stream_s *dy_stream_t;        /* a hipStream_t */
typedef struct dy_event_s *dy_event_t;          /* a hipEvent_t */
typedef struct dy_broadphase_s *dy_broadphase_t;
typedef struct dy_narrowphase_s *dy_narrowphase_t;
typedef struct dy_xpbd_s *dy_xpbd_t;
typedef struct dy_graph_s *dy_graph_t;          /* an instantiated hipGraph */

/* ---- runtime (compute/include/compute/{device,stream,buffer}.hpp) -------- */
const char *dy_status_string(int status);
int dy_version(int *major, int *minor);
/* Device::default_gpu / name / is_gpu / max_compute_units / global_mem_bytes (device.hpp:12-37) */
int dy_device_count(int *count);
int dy_device_get_info(int device, dy_device_info *out);
int dy_set_device(int device);
/* Stream(const sycl::device&) in-order queue; Stream::wait (stream.hpp:18-30) */
int dy_stream_create(dy_stream_t *out);
int dy_stream_destroy(dy_stream_t s);
int dy_stream_synchronize(dy_stream_t s);
/* Buffer<T>(Stream&, n) / ~Buffer (buffer.hpp:15-19) */
int dy_malloc(void **ptr, size_t bytes);
int dy_free(void *ptr);
/* Buffer::upload (async), Buffer::download (blocking), queue().memcpy/memset (buffer.hpp:42-49) */
int dy_memcpy_h2d_async(void *dst, const void *src, size_t bytes, dy_stream_t s);
int dy_memcpy_d2h(void *dst, const void *src, size_t bytes, dy_stream_t s);
int dy_memcpy_d2d_async(void *dst, const void *src, size_t bytes, dy_stream_t s);
/* queue.memcpy(dst, src, bytes) (compute/include/compute/buffer.hpp:45-53):
 * direction taken from the pointers (host pageable, pinned or device). */
int dy_memcpy_async(void *dst, const void *src, size_t bytes, dy_stream_t s);
int dy_memset_async(void *dst, int value, size_t bytes, dy_stream_t s);
/* Event (stream.hpp:6-14), plus timing */
int dy_event_create(dy_event_t *out);
int dy_event_destroy(dy_event_t e);
int dy_event_record(dy_event_t e, dy_stream_t s);
int dy_event_synchronize(dy_event_t e);
int dy_event_elapsed_ms(dy_event_t start, dy_event_t stop, float *ms);
/* HIP graph capture of a stream (no reference counterpart: the SYCL queue
 * records nothing).  Calls issued on s between begin and end are recorded,
 * not run; end instantiates them.  Launching replays them in order on a
 * stream; destroy frees the graph. */
int dy_graph_capture_begin(dy_stream_t s);
int dy_graph_capture_end(dy_stream_t s, dy_graph_t *out);
int dy_graph_launch(dy_graph_t g, dy_stream_t s);
int dy_graph_destroy(dy_graph_t g);

/* ---- compute primitives ------------------------------------------------- */
/* sort_by_key / sort (compute/include/compute/sort.hpp:29,89): ascending, stable.
 * Unlike the reference, n need not be a power of two.  Async. */
int dy_sort_by_key_u64_u32(dy_stream_t s, uint64_t *keys, uint32_t *vals, size_t n);
int dy_sort_by_key_u32_u32(dy_stream_t s, uint32_t *keys, uint32_t *vals, size_t n);
int dy_sort_u32(dy_stream_t s, uint32_t *keys, size_t n);
int dy_sort_u64(dy_stream_t s, uint64_t *keys, size_t n);
/* reduce_sum (compute/include/compute/reduction.hpp:22-62): deterministic
 * fixed-order tree; blocking, result to host. */
int dy_reduce_sum_f32(dy_stream_t s, const float *d_data, size_t n, float *out);
/* Exclusive prefix sum of u32 with the total written to d_total (may be NULL). Async. */
int dy_scan_exclusive_u32(dy_stream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n,
                          uint32_t *d_total);

/* ---- host set-up --------------------------------------------------------- */
/* BodyStore::add dynamics (core/src/body_store.cpp:121-139, mat3.hpp:61-79) for n
 * bodies: mass[n], inertia[9n] row-major, flags_in[n] -> inv_mass[n], iI[6n]
 * (xx yy zz xy xz yz interleaved per body), flags_out[n].  Host memory, no GPU. */
int dy_body_dynamics_host(uint32_t n, const float *mass, const float *inertia,
                          const uint32_t *flags_in, float *inv_mass, float *iI,
                          uint32_t *flags_out);

/* ---- stages --------------------------------------------------------------- */
/* integrate (core/include/core/integrator.hpp:25, core/src/integrator.cpp:7-71).
 * All substeps in one launch (each body's substeps depend only on itself). */
int dy_integrate(dy_stream_t s, const dy_body_view *bodies, float gx, float gy, float gz,
                 float dt, int n_substeps);

/* Broadphase(Stream&, max_bodies, max_pairs) (core/include/core/broadphase.hpp:57) */
int dy_broadphase_create(dy_stream_t s, uint32_t max_bodies, uint32_t max_pairs,
                         dy_broadphase_t *out);
int dy_broadphase_destroy(dy_broadphase_t bp);
/* build_and_query (broadphase.hpp:62-65): Morton sort, Karras tree, fp16 refit,
 * and the candidate pairs, written in ascending (a, b) order.  No host sync.
 * Pairs beyond max_pairs are dropped (the first max_pairs in (a, b) order are
 * kept); the device count holds the true number. */
int dy_broadphase_build_and_query(dy_broadphase_t bp, dy_stream_t s, const dy_body_view *bodies,
                                  const dy_shape_view *shapes, const dy_aabb *scene_bounds);
/* pairs_ptr / count_ptr / max_pairs (broadphase.hpp:68-70) */
int dy_broadphase_pairs(dy_broadphase_t bp, const dy_contact_pair **pairs, const uint32_t **d_count,
                        uint32_t *max_pairs);
/* sort_pairs (broadphase.hpp:75): the pairs are already in canonical order, so
 * this only validates its argument. */
int dy_broadphase_sort_pairs(dy_broadphase_t bp, dy_stream_t s, uint32_t n_pairs);
/* download_count (broadphase.hpp:78): blocking; the true count. */
int dy_broadphase_download_count(dy_broadphase_t bp, dy_stream_t s, uint32_t *out);
/* Enable (1) or disable (0) kernel timing; enabling resets the totals. */
int dy_broadphase_set_timing(dy_broadphase_t bp, int enable);
/* Totals of the timed builds; blocking (waits for the last timed build). */
int dy_broadphase_get_timing(dy_broadphase_t bp, dy_broadphase_timing *out);
/* bvh_view (broadphase.hpp:82) */
int dy_broadphase_bvh_view(dy_broadphase_t bp, uint32_t n_active, dy_bvh_view *out);

/* Narrowphase(Stream&, max_contacts) (core/include/core/narrowphase.hpp:29) */
int dy_narrowphase_create(dy_stream_t s, uint32_t max_contacts, dy_narrowphase_t *out);
int dy_narrowphase_destroy(dy_narrowphase_t np);
/* run(s, pairs, n_pairs, bodies, shapes) (narrowphase.hpp:33-37): one contact per
 * colliding pair, stored in pair order; no host sync.  Contacts beyond capacity
 * are dropped in pair order; the true count is kept. */
int dy_narrowphase_run(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs,
                       uint32_t n_pairs, const dy_body_view *bodies, const dy_shape_view *shapes);
/* run(s, pairs, d_n_pairs, ...) (narrowphase.hpp:42-46): the pair count stays on
 * the device; max_pairs bounds the launch (entries beyond *d_n_pairs are ignored). */
int dy_narrowphase_run_device_count(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs,
                                    const uint32_t *d_n_pairs, uint32_t max_pairs,
                                    const dy_body_view *bodies, const dy_shape_view *shapes);
/* contacts() (narrowphase.hpp:48) */
int dy_narrowphase_contacts(dy_narrowphase_t np, dy_contact_view *out);
/* download_count (narrowphase.hpp:49): blocking; stored count.  Returns
 * DY_ERR_UNSUPPORTED if the last run met a hull or mesh shape. */
int dy_narrowphase_download_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out);
/* last_contact_count (narrowphase.hpp:57): blocking here; the true count. */
int dy_narrowphase_last_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out);
/* Feature words (no reference counterpart; core/src/narrowphase.cpp:153-305
 * computes their parts and keeps only the contact).  set_features(np, 1)
 * makes later runs also store, per stored contact in contact order, a word:
 *   bits 0-2   kind: 0 box-box with A's face as the reference face, 1 box-box
 *              with B's face, 2 box-box edge-edge, 3 sphere-sphere, 4 sphere-box
 *              with the centre outside (closest point), 5 sphere-box with the
 *              centre inside
 *   bits 3-4   the reference-face axis (kinds 0, 1), A's edge axis (2), the face
 *              of least penetration (5); else 0
 *   bits 5-6   B's edge axis (kind 2); else 0
 *   bits 8-15  kinds 0, 1: the incident box's vertices k = 0..7 (signs of the
 *              half extents x, y, z by bits 0, 1, 2 of k) that passed the depth
 *              and both side tests; the first four in k order form the manifold
 *   bits 16-19 box-box: the SAT axis of least overlap before the reference-face
 *              choice (0-2 A's faces, 3-5 B's faces, 6 + 3i + j the edge pair
 *              A_i x B_j)
 * features() gives the device array (DY_ERR_INVALID_ARG while disabled). */
int dy_narrowphase_set_features(dy_narrowphase_t np, int enable);
int dy_narrowphase_features(dy_narrowphase_t np, const uint32_t **d_features);
/* Allocation-free narrowphase.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same pass and results
 * as dy_narrowphase_run_device_count over the first min(*d_n_pairs, max_pairs)
 * pairs (all max_pairs when d_n_pairs is NULL), written into the caller's
 * store: store's arrays and capacity (store->n is not read), d_counts =
 * DY_NARROWPHASE_COUNT_WORDS device words, on return [0] the stored count,
 * [1] the true count, [2] nonzero if a hull or mesh shape was met, [3]
 * nonzero on an internal error ([4] is working space); d_features (may be
 * NULL) receives a feature word per stored contact.  A contact view with
 * n = d_counts feeds dy_xpbd_solve.  Calls no hipMalloc/hipFree: its only
 * working memory is a fixed 512 KiB device array per GPU, so max_pairs is at
 * most 16 777 216 (larger: DY_ERR_INVALID_ARG).  Calls that overlap on one GPU
 * (other streams, other threads) share that array and race; calls in one
 * stream's order do not.  Asynchronous. */
#define DY_NARROWPHASE_COUNT_WORDS 5
int dy_narrowphase_run_noalloc(dy_stream_t s, const dy_contact_pair *d_pairs, const uint32_t *d_n_pairs,
                               uint32_t max_pairs, const dy_body_view *bodies, const dy_shape_view *shapes,
                               const dy_contact_view *store, uint32_t *d_counts, uint32_t *d_features);

/* XpbdSolver(Stream&, n_iters, max_contacts, friction) (core/include/core/xpbd_solver.hpp:31-34) */
int dy_xpbd_create(dy_stream_t s, int n_iters, uint32_t max_contacts, float friction, dy_xpbd_t *out);
int dy_xpbd_destroy(dy_xpbd_t x);
/* set_force_serial (xpbd_solver.hpp:48) */
int dy_xpbd_set_force_serial(dy_xpbd_t x, int force_serial);
/* solve(s, contacts, joints, bodies, dt) (core/include/core/solver.hpp:17-21,
 * core/src/xpbd_solver.cpp:590-733).  Constraint k is contact k (k < nc) or
 * joint k - nc; joints may be NULL (= JointView{}).  <= 256 constraints or
 * force_serial: the reference's serial order (run_serial: joint motors in
 * joint order, then per iteration contacts, joints, renormalisation),
 * executed in parallel over dependency levels; otherwise the reference's
 * Jones-Plassmann colours over contacts and joints, motors then constraints
 * swept in colour order, falling back to the serial order where the
 * reference does (a constraint needing a 65th colour, or more than 256
 * rounds).  The contact count is read on the device (*contacts->n, clamped
 * to max_contacts; a larger count is an error).  Joints are validated on the
 * device: a body index out of range, parent == child or a type above 3 makes
 * that solve run without joints and the solver report DY_ERR_INVALID_ARG.
 * Asynchronous on the stream (the launch-per-round path, DYPHUR_SOLVER_PATH=1,
 * synchronises; a solve with more joints than any before it synchronises the
 * device to grow its buffers).  Errors of a solve are returned by a later
 * solve or by the blocking queries below. */
int dy_xpbd_solve(dy_xpbd_t x, dy_stream_t s, const dy_contact_view *contacts, const dy_joint_view *joints,
                  const dy_body_view *bodies, float dt);
/* Statistics of the last solve; blocking (waits for that solve). */
int dy_xpbd_last_stats(dy_xpbd_t x, dy_solve_stats *out);
/* Enable (1) or disable (0) kernel timing; enabling resets the totals. */
int dy_xpbd_set_timing(dy_xpbd_t x, int enable);
/* Totals of the timed solves; blocking (waits for the stream's last solve). */
int dy_xpbd_get_timing(dy_xpbd_t x, dy_stream_t s, dy_solver_timing *out);
/* Class (colour or level) of each contact of the last solve; blocking. */
int dy_xpbd_download_classes(dy_xpbd_t x, dy_stream_t s, uint32_t *out, uint32_t n);

/* ---- ray queries ------------------------------------------------------------
 * Ray / RayHit (core/include/core/ray_query.hpp:9-18): direction need not be
 * normalised, t is in direction units; a hit is recorded only for t < t_max. */
typedef struct dy_ray { float ox, oy, oz, dx, dy, dz, t_max; } dy_ray;
typedef struct dy_ray_hit { float t; uint32_t body_idx; } dy_ray_hit;
/* ray_query(s, bvh, bodies, shapes, d_rays, d_hits, n_rays)
 * (core/include/core/ray_query.hpp:35-41, core/src/ray_query.cpp:126-195):
 * closest exact hit (oriented box, sphere) over the BVH of the last
 * build_and_query (dy_broadphase_bvh_view); body_idx 0xFFFFFFFF and t = t_max
 * when nothing is hit.  Device arrays; asynchronous. */
int dy_ray_query(dy_stream_t s, const dy_bvh_view *bvh, const dy_body_view *bodies, const dy_shape_view *shapes,
                 const dy_ray *d_rays, dy_ray_hit *d_hits, uint32_t n_rays);

/* ---- batched scenes -------------------------------------------------------
 * Many independent scenes stepped by one launch: the drop-in frame
 * (examples/stress_test_blocks/main.cpp:264-273) of every scene, one workgroup
 * per scene.  The reference has no batched form; each scene's results equal
 * those of a stand-alone world over the same bodies (pairs in canonical
 * order, contacts in pair order, reference-order or coloured solve by the
 * reference's rules), with scene-local body indices.  Scene s owns bodies
 * [body_offset[s], body_offset[s+1]) of the body view passed to each step;
 * body shape indices address the shape view passed with it.  At most 2048
 * bodies per scene. */
typedef struct dy_batch_s *dy_batch_t;

/* Per-scene statistics of the scene's last substep, and running totals. */
typedef struct dy_batch_scene_stats {
    uint32_t n_pairs;      /* true candidate-pair count (may exceed the scene's pair capacity) */
    uint32_t n_contacts;   /* true contact count (may exceed the scene's contact capacity) */
    uint32_t n_classes;    /* colours (coloured path) or dependency levels (reference order) */
    uint32_t used_serial;  /* 1: reference contact order (<= 256 contacts, forced, or fallback) */
    uint32_t rounds;       /* class-assignment rounds of the last solve */
    uint32_t overflow;     /* 1: a body needed a 65th colour */
    uint32_t error;        /* OR of: 1 hull/mesh shape met, 2 internal check failed */
    uint32_t substeps;     /* substeps run since creation */
    uint32_t sum_pairs;    /* totals over those substeps: true pairs, */
    uint32_t sum_contacts; /* stored contacts, */
    uint32_t sum_dyn_endpoints;  /* contact endpoints on dynamic bodies, */
    uint32_t solves;       /* substeps that solved at least one contact */
    uint32_t grid_substeps;  /* substeps whose pairs came from the scene's x-z grid rather than every pair */
} dy_batch_scene_stats;

/* Device outputs of the last substep, concatenated over scenes. */
typedef struct dy_batch_layout {
    const dy_contact_pair *pairs;    /* scene s: pair_offset[s] .. + min(n_pairs, capacity) */
    dy_contact_view contacts;        /* scene s: contact_offset[s] ..; contacts.n is NULL */
    const uint32_t *body_offset;     /* host arrays of n_scenes + 1 entries */
    const uint32_t *pair_offset;
    const uint32_t *contact_offset;
    uint32_t n_scenes;
} dy_batch_layout;

/* body_offset: host [n_scenes + 1], body_offset[0] = 0, strictly increasing;
 * max_pairs / max_contacts: host [n_scenes], per-scene capacities as
 * Broadphase(s, n, max_pairs) / Narrowphase(s, max_contacts) take them;
 * n_iters / friction as XpbdSolver.  Blocking (uploads the layout). */
int dy_batch_create(dy_stream_t s, uint32_t n_scenes, const uint32_t *body_offset, const uint32_t *max_pairs,
                    const uint32_t *max_contacts, int n_iters, float friction, dy_batch_t *out);
int dy_batch_destroy(dy_batch_t b);
/* set_force_serial for every scene (xpbd_solver.hpp:48). */
int dy_batch_set_force_serial(dy_batch_t b, int force_serial);
/* Launch shape of the step kernel: workgroup threads (64, 128 or 256) and the
 * largest scene (bodies, ground included) stepped by one wavefront; larger
 * scenes get a whole workgroup, smaller ones are packed threads/64 to a
 * workgroup.  Default 256 / 128.  Results do not depend on the shape. */
int dy_batch_set_shape(dy_batch_t b, uint32_t threads, uint32_t wave_scene_max);
int dy_batch_get_shape(dy_batch_t b, uint32_t *threads, uint32_t *wave_scene_max);
/* LDS one workgroup of the step may take, in bytes (0: the device's maximum,
 * the default: 160 KB on gfx950; values below 64 KiB read as 64 KiB).  Scenes
 * whose state (112 B per body) fits 64 KiB share launches; larger ones that
 * fit the limit run one per workgroup in a launch of their own; the rest keep
 * their state in global memory.  The reverse sweep of a dy_diff_t created
 * afterwards is planned with the same limit.  Results do not depend on it.
 * Blocking when the limit changes (waits for the device, then re-plans).
 * No counterpart in the reference (its scenes never leave global memory). */
int dy_batch_set_lds_limit(dy_batch_t b, uint32_t bytes);
/* n_frames frames of every scene in ONE launch; each frame is n_substeps
 * substeps of integrate(dt / n_substeps, 1) -> pairs -> narrowphase ->
 * solve(dt / n_substeps).  Asynchronous; issues no memset or copy, so it can be
 * captured in a HIP graph. */
int dy_batch_step(dy_batch_t b, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes,
                  float gx, float gy, float gz, float dt, int n_substeps, int n_frames);
/* Times trial_frames frames at each of six launch shapes (threads /
 * wave_scene_max: 64/-, 128/64, 128/128, 256/64, 256/128, 256/256) from the
 * current state, keeps the fastest, and restores the body state and
 * statistics.  Blocking.  chosen_* (may be NULL) receive the shape; ms_out
 * (may be NULL) the six times in ms. */
int dy_batch_autotune(dy_batch_t b, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes,
                      float gx, float gy, float gz, float dt, int n_substeps, int trial_frames,
                      uint32_t *chosen_threads, uint32_t *chosen_wave_max, float *ms_out);
/* Per-scene statistics, n_scenes entries; blocking.  Returns
 * DY_ERR_UNSUPPORTED when a scene met a hull or mesh shape since creation and
 * DY_ERR_INTERNAL on any other device error flag (out is filled either way). */
int dy_batch_stats(dy_batch_t b, dy_stream_t s, dy_batch_scene_stats *out);
int dy_batch_layout_get(dy_batch_t b, dy_batch_layout *out);
/* Per-stage timing of the step kernel (one clock read per stage boundary by
 * each workgroup).  Enabling clears the totals (asynchronous on s); get
 * copies 5 counters per scene (integrate + leaf boxes, pairs, narrowphase,
 * class assignment, sweep) in ticks of tick_hz, blocking. */
int dy_batch_set_timing(dy_batch_t b, dy_stream_t s, int enable);
int dy_batch_get_timing(dy_batch_t b, dy_stream_t s, uint64_t *ticks, double *tick_hz);

/* ---- differentiable batched form ------------------------------------------
 * No reference counterpart.  Reverse-mode derivatives of a batch's frames
 * (n_substeps substeps each) with respect to the initial state and the
 * inverse masses.  Forward: the batch's frames, the mutable state saved every
 * checkpoint_every frames.  Backward: per checkpoint segment, newest first,
 * the segment's substeps are recomputed with a record (the tape) of every
 * intermediate the adjoints need, then reversed: hand-written adjoints of
 * the integrator, the contact projection and the quaternion
 * renormalisation, and the contact geometry's Jacobian with respect to the
 * two poses.  The contact set, pair list and solve order are those of the
 * forward (piecewise-smooth derivative).  A segment longer than a tape is
 * replayed once to save a state per tape-sized chunk, then recomputed and
 * reversed chunk by chunk.  Statistics of the batch count the recomputed
 * substeps too. */
typedef struct dy_diff_s *dy_diff_t;
/* A tape step holds, per scene, tape_contacts_per_body x bodies contacts
 * (<= 0: the scene's contact capacity); a step with more contacts makes the
 * backward fail with DY_ERR_INVALID_ARG.  Each of the two tapes takes at most
 * tape_bytes (0: 24 GiB), in whole frames; DY_ERR_INVALID_ARG if one frame
 * does not fit.  When the records of all ceil(max_frames / checkpoint_every)
 * segments fit 2 x tape_bytes, the forward records every segment instead and
 * the backward recomputes nothing (same gradient bits).  Allocations are
 * done here (blocking). */
int dy_diff_create(dy_batch_t b, uint32_t max_frames, uint32_t checkpoint_every, uint32_t n_substeps,
                   float tape_contacts_per_body, uint64_t tape_bytes, dy_diff_t *out);
int dy_diff_destroy(dy_diff_t d);
/* Substeps one tape holds, and the bytes one recorded substep of every scene
 * takes (may be NULL). */
int dy_diff_tape_steps(dy_diff_t d, uint32_t *steps, uint64_t *bytes_per_step);
/* n_frames (<= max_frames) frames of every scene from the state in bodies,
 * which ends in the final state.  Asynchronous. */
int dy_diff_forward(dy_diff_t d, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes,
                    float gx, float gy, float gz, float dt, int n_frames);
/* After dy_diff_forward: adj_final holds the loss's adjoint of the final
 * state (its 13 mutable arrays: pos, rot, vel, ang); adj_initial receives the
 * adjoint of the initial state in the same 13 arrays and the adjoint of the
 * inverse masses in its inv_mass array (the two views may share arrays).
 * Other fields of both views are unused.  Asynchronous, except with a
 * reduced tape (it then waits to report an overfull step). */
int dy_diff_backward(dy_diff_t d, dy_stream_t s, const dy_body_view *bodies, const dy_shape_view *shapes,
                     float gx, float gy, float gz, float dt, const dy_body_view *adj_final,
                     dy_body_view *adj_initial);

/* FNV-1a-64 over pos xyz, rot wxyz of all bodies
 * (examples/stress_test_blocks/main.cpp:34-41,301-311); blocking. */
int dy_state_hash_fnv1a64(dy_stream_t s, const dy_body_view *bodies, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* DYPHUR_HIP_H */
