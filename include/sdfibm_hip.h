/*
 * sdfibm_hip.h -- C ABI of the MI355X (gfx950) solid-fluid coupling step.
 *
 * One call of sdfibm_interact() is the reference's SolidCloud::interact(time, dt)
 * (src/solidcloud.cpp:435-464): zero the solid fields, for every solid in index
 * order enumerate the cells it touches (src/cellenumerator.cpp:6-78), compute
 * each cell's solid volume fraction (src/geometrictools.cpp:6-116), add the
 * direct-forcing terms and reduce the fluid force and torque on the solid
 * (src/solidcloud.cpp:361-433), then clip As to at most 1
 * (src/solidcloud.cpp:564-570).
 *
 * The mesh, U, As, Fs, Ts and Ct live on the device.  Plain C: pointers, sizes
 * and status codes; no exceptions cross this boundary.  Every entry point
 * returns SDFIBM_OK (0) or a negative sdfibm_status; sdfibm_last_error() gives
 * the message of the last failure on a context.
 *
 * Threading: one context per host thread; calls on one context are not
 * reentrant.  Ownership: the caller owns every host array it passes; the
 * context owns its device memory.
 */
#ifndef SDFIBM_HIP_H
#define SDFIBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDFIBM_ABI_VERSION 1

typedef enum {
    SDFIBM_OK = 0,
    SDFIBM_ERR_ARG = -1,       /* bad argument (null pointer, size, dt <= 0, NaN state) */
    SDFIBM_ERR_HIP = -2,       /* a HIP runtime call failed (message has the HIP error) */
    SDFIBM_ERR_NOMEM = -3,     /* device allocation failed */
    SDFIBM_ERR_SHAPE = -4,     /* unknown shape id or invalid shape parameters */
    SDFIBM_ERR_TILE = -5,      /* a solid's cell tile grew beyond 2^22 vertices during run_steps' graph
                                  replays (such a solid goes the planes' path instead -- at set_solids,
                                  or at the interact that finds it grown; <= 32 such plus planes) */
    SDFIBM_ERR_STATE = -6,     /* call out of order (e.g. interact before set_solids) */
    SDFIBM_ERR_MESH = -7,      /* unsupported or inconsistent mesh description */
    SDFIBM_ERR_CAPACITY = -8,  /* an internal buffer was too small for this step */
    SDFIBM_ERR_NODEVICE = -9   /* no HIP device available */
} sdfibm_status;

/* Shape ids.  Names and parameters follow the reference's shape plugins
 * (src/libshape/ headers, registered by type name in shapefactory.h:11-39). */
typedef enum {
    SDFIBM_SHAPE_SPHERE = 0,         /* sphere.h:10-47        prm: radius, com(3)             */
    SDFIBM_SHAPE_ELLIPSOID = 1,      /* ellipsoid.h:10-55     prm: radiusa, radiusb, radiusc  */
    SDFIBM_SHAPE_CIRCLE = 2,         /* circle.h:10-47        prm: radius, com(3)             */
    SDFIBM_SHAPE_ELLIPSE = 3,        /* ellipse.h:10-56       prm: radiusa, radiusb, -, com(3) */
    SDFIBM_SHAPE_BOX = 4,            /* box.h:10-59           prm: radiusa,b,c, com(3)        */
    SDFIBM_SHAPE_RECTANGLE = 5,      /* rectangle.h:10-60     prm: radiusa, radiusb, -, com(3) */
    SDFIBM_SHAPE_PLANE = 6,          /* plane.h:10-30         no parameters; inside: local y<0 */
    SDFIBM_SHAPE_CIRCLE_TAIL = 7,    /* circle_tail.h:10-63   prm: radius, ratio, thickness, com(3) */
    SDFIBM_SHAPE_CIRCLE_TWOTAIL = 8, /* circle_twotail.h:10-68 prm: radius, ratio, thickness, com(3) */
    SDFIBM_SHAPE_CSG = 9             /* prm[0]: index of its shape in the table of sdfibm_set_csg_shapes */
} sdfibm_shape_id;

/* CSG shapes: up to SDFIBM_CSG_MAX_PRIMS primitives in the solid's local frame,
 * combined left to right with the reference's CSG operators (src/libshape/sdf/sdf.h:133-142:
 * union U = min / or, difference D(a, b) = max(a, -b) / a and not b, intersection
 * I = max / and), the signed distance filtered at the end like every shape
 * (sdf.h:147-150).  Primitives: box and ellipsoid as sdf.h:43-57 and 72-83 (the
 * ellipsoid's first-order distance), sphere, and a finite cylinder (distance: the
 * rectangle formula of sdf.h:29-37 in radius/axial coordinates).  The reference has
 * the operators and primitives but no 3-D CSG shape plugin; this one is the
 * library's.  Like a shape plugin it declares its mass properties (IShape::m_volume,
 * m_moi, m_radiusB, ishape.h:32-38), used by evolve and the contacts. */
#define SDFIBM_CSG_MAX_PRIMS 8
typedef enum { SDFIBM_CSG_BOX = 0, SDFIBM_CSG_ELLIPSOID = 1, SDFIBM_CSG_CYLINDER = 2, SDFIBM_CSG_SPHERE = 3 } sdfibm_csg_kind;
typedef enum { SDFIBM_CSG_UNION = 0, SDFIBM_CSG_DIFFERENCE = 1, SDFIBM_CSG_INTERSECTION = 2 } sdfibm_csg_op;
typedef struct {
    int32_t kind;      /* sdfibm_csg_kind */
    int32_t op;        /* sdfibm_csg_op combining it with the primitives before it (ignored for the first) */
    int32_t axis;      /* cylinder axis in the solid's frame: 0, 1 or 2 */
    int32_t reserved;
    double center[3];  /* in the solid's frame */
    double size[3];    /* box: half sides; ellipsoid: semi-axes; cylinder: radius, half length; sphere: radius */
} sdfibm_csg_prim;     /* 64 bytes */
typedef struct {
    int32_t n_prims;   /* 1 .. SDFIBM_CSG_MAX_PRIMS */
    int32_t reserved;
    double volume;     /* > 0 */
    double moi[3];     /* principal moments per unit density about the solid's origin, > 0 */
    double radius_b;   /* bounding radius for the contact search, >= 0 */
    sdfibm_csg_prim prim[SDFIBM_CSG_MAX_PRIMS];
} sdfibm_csg_shape;    /* 560 bytes */

/* One solid: shape + state.  Replaces Solid (src/solid.h:14-50) as far as the
 * coupling step reads it: centre, orientation, velocity, omega
 * (solid.h:19-26) and the shape it points at (solid.h:48).  160 bytes. */
typedef struct {
    int32_t shape;     /* sdfibm_shape_id */
    int32_t flags;     /* reserved, 0 */
    double prm[6];     /* shape parameters, solidDict units */
    double center[3];
    double q[4];       /* orientation quaternion (w, x, y, z); world2local = conj(q) (p - center) q */
    double vel[3];
    double omega[3];
} sdfibm_solid;

typedef enum {
    SDFIBM_MESH_RECTILINEAR = 0, /* points = xs x ys x zs tensor product; geometry in closed form */
    SDFIBM_MESH_STRUCTURED = 1,  /* one hex block with arbitrary point positions; geometry given */
    SDFIBM_MESH_POLY = 2         /* a general polyMesh (sdfibm_create_polymesh) */
} sdfibm_mesh_kind;

/* A logically structured hex mesh (one blockMesh block): points numbered
 * i-fastest then j then k, cells likewise, cell points and faces in OpenFOAM's
 * hex-model order.  Replaces MeshInfo (src/meshinfo.h:11-29): cellCells and
 * cellPoints are implicit in (i, j, k); points, cell centres, volumes, face
 * centres and face area vectors are the arrays below. */
typedef struct {
    int32_t kind;             /* sdfibm_mesh_kind */
    int32_t nx, ny, nz;       /* cells per direction */
    /* RECTILINEAR */
    const double* xs;         /* [nx+1] */
    const double* ys;         /* [ny+1] */
    const double* zs;         /* [nz+1] */
    /* STRUCTURED (all AoS xyz, FP64, as OpenFOAM stores them) */
    const double* points;         /* [(nx+1)(ny+1)(nz+1) * 3] */
    const double* cell_centres;   /* [nx*ny*nz * 3]  (mesh.C())  */
    const double* cell_volumes;   /* [nx*ny*nz]      (mesh.V())  */
    const double* face_centres[3];/* x-faces [(nx+1)*ny*nz*3], y-faces [nx*(ny+1)*nz*3], z-faces [nx*ny*(nz+1)*3] */
    const double* face_areas[3];  /* same layout; area vectors (mesh.Sf()), sign ignored */
} sdfibm_mesh_desc;

/* A general polyhedral mesh: OpenFOAM's polyMesh addressing (points, faces,
 * owner, neighbour) and the geometry OpenFOAM derives from it.  Replaces
 * MeshInfo (src/meshinfo.h:11-29) for any mesh the reference runs on: several
 * blockMesh blocks, O-grids, snappyHexMesh output.  The library derives cells()
 * (each cell's owned faces ascending, then its neighbour faces ascending, as
 * primitiveMesh::calcCells lists them), cellPoints (ascending point labels) and
 * cellCells from the addressing.  The four geometry arrays may be NULL: they are
 * then computed as primitiveMesh does (face centres and area vectors from the
 * triangles fanned about the face's point average; cell centres and volumes from
 * the face pyramids about the average of the cell's face centres). */
typedef struct {
    int64_t n_points, n_faces, n_internal_faces, n_cells;
    const double* points;         /* [n_points*3]                                    (mesh.points()) */
    const int64_t* face_offsets;  /* [n_faces+1]: face f's points are                              */
    const int32_t* face_points;   /*   face_points[face_offsets[f] .. face_offsets[f+1])  (mesh.faces()) */
    const int32_t* owner;         /* [n_faces]                                       (mesh.faceOwner()) */
    const int32_t* neighbour;     /* [n_internal_faces]                              (mesh.faceNeighbour()) */
    const double* cell_centres;   /* [n_cells*3] or NULL                             (mesh.C()) */
    const double* cell_volumes;   /* [n_cells] or NULL                               (mesh.V()) */
    const double* face_centres;   /* [n_faces*3] or NULL                             (mesh.Cf()) */
    const double* face_areas;     /* [n_faces*3] or NULL                             (mesh.Sf()) */
} sdfibm_polymesh_desc;

typedef enum { SDFIBM_FP32 = 0, SDFIBM_FP64 = 1 } sdfibm_precision;

typedef struct {
    int32_t two_d;        /* meta.on_twod: cell apex z set to 0 (geometrictools.cpp:57-58) */
    int32_t precision;    /* sdfibm_precision of the resident fields and the fraction arithmetic */
    double rho_f;         /* fluid density multiplying force/torque; 0 when on_fluid = 0
                             (solidcloud.cpp:247-251, 424-425) */
    int64_t reserved[4];
} sdfibm_config;

typedef struct sdfibm_ctx sdfibm_ctx;

/* Create a context on the current HIP device: copies the mesh to the device
 * once and allocates U, As, Fs, Ts, Ct.  Replaces the SolidCloud constructor's
 * field binding (solidcloud.cpp:209-217) and MeshInfo (meshinfo.h:22-29). */
int sdfibm_create(const sdfibm_mesh_desc* mesh, const sdfibm_config* cfg, sdfibm_ctx** out);

/* The same for a general polyhedral mesh (the constructor's MeshInfo binding,
 * meshinfo.h:22-29, for any OpenFOAM mesh).  Every call below works on such a
 * context except sdfibm_run_steps, sdfibm_interact_steps, sdfibm_autotune and the
 * surrogate (SDFIBM_ERR_STATE): its step reads three counts back to size its
 * buffers, so it is not captured as a graph.  sdfibm_last_timing: [0] the
 * (solid, cell) pairs, their labels and force/torque, [1] the per-cell lists,
 * [2] the field pass, [3] total. */
int sdfibm_create_polymesh(const sdfibm_polymesh_desc* mesh, const sdfibm_config* cfg, sdfibm_ctx** out);
int sdfibm_destroy(sdfibm_ctx* ctx);

/* Upload the solids (index order = the reference's m_solids order,
 * solidcloud.h:49).  Validates shape ids and parameters.  Solids of any size:
 * tiles beyond the on-chip tile pass (more than 63 cells along x or 1024 vertex
 * rows; at most 16384 such solids) take the big-tile pass; beyond 2^22 tile
 * vertices a solid takes the planes' path (<= 32 such plus planes). */
int sdfibm_set_solids(sdfibm_ctx* ctx, const sdfibm_solid* solids, int64_t n);

/* The CSG shape table the context's CSG solids refer to (copied).  The solids
 * must be set again afterwards (SDFIBM_ERR_STATE until then). */
int sdfibm_set_csg_shapes(sdfibm_ctx* ctx, const sdfibm_csg_shape* shapes, int64_t n);

/* Fluid velocity U: host AoS FP64 [N*3] (a volVectorField's internal field),
 * converted to the resident SoA layout in the context's precision. */
int sdfibm_set_velocity(sdfibm_ctx* ctx, const double* U_aos);

/* Fluid velocity of cells [first, first + count), host AoS FP64 [count*3]: what
 * sdfibm_set_velocity does for a range of cells, without allocating device
 * memory.  The values pass through a fixed staging area of
 * SDFIBM_NOALLOC_MAX_CELLS cells (24 MiB) that each GPU the library runs on holds
 * until the process ends; a larger count is SDFIBM_ERR_ARG (upload in pieces).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (solver loops and serving paths that forbid hipMalloc).  All contexts
 * on one GPU share the staging area: calls that overlap in time on one GPU (from
 * several host threads) race; the library does not serialise them. */
#define SDFIBM_NOALLOC_MAX_CELLS 1048576
int sdfibm_set_velocity_cells_noalloc(sdfibm_ctx* ctx, int64_t first, int64_t count, const double* U_aos);

/* SolidCloud::interact(time, dt) (solidcloud.cpp:435-464).  force_torque, if
 * not NULL, receives [n_solids * 6] FP64 (fx fy fz tx ty tz) per solid, the
 * values the reference stores with setFluidForceAndTorque (solid.h:136-140).
 * Fields stay on the device. */
int sdfibm_interact(sdfibm_ctx* ctx, double time, double dt, double* force_torque);

/* The step after interact in the reference's solver loop: U = U - Fs*dt over
 * the resident fields (src/main.cpp:70). */
int sdfibm_update_velocity(sdfibm_ctx* ctx, double dt);

/* SolidCloud::fixInternal (src/solidcloud.cpp:288-301): in every cell with
 * Ct >= 4 (wholly inside solid Ct-4, from the last interact) U becomes that
 * solid's rigid velocity v + omega x (cc - centre), with the solid state last
 * passed to sdfibm_set_solids. */
int sdfibm_fix_internal(sdfibm_ctx* ctx);

/* Per-solid material and motion constraint of the rigid-body update.  Replaces
 * Solid::setMaterial (solid.h:94-103: the material's density) and the
 * Motion01Mask plugin (libmotion/motion01mask.h:15-64). */
typedef struct {
    double rho;           /* solid density, > 0 */
    int32_t motion_mask;  /* bits 0..5 = vx vy vz wx wy wz free, i.e. Motion01Mask "b<vx><vy><vz><wx><wy><wz>"
                             with bit k = character k+1; -1 = no motion constraint */
    int32_t reserved;
} sdfibm_dynamics;

/* Set the dynamics of every solid (one record per solid of the last
 * set_solids); resets the first-step state of the fluid force mix. */
int sdfibm_set_dynamics(sdfibm_ctx* ctx, const sdfibm_dynamics* dyn, int64_t n);

/* SolidCloud::evolve(time, dt) (solidcloud.cpp:521-562) on the device: 20
 * sub-steps (1 for a single solid) of AB2 fluid force from the last interact
 * (solid.h:164-174), buoyant gravity ((rho_s - rho_f)/rho_s g,
 * solidcloud.cpp:466-475) and Solid::move (solid.h:182-202: velocity, world
 * inverse inertia, motion mask, trapezoidal position, quaternion update and
 * normalisation).  gravity: [3] or NULL.  The solids stay on the device; the
 * next interact and fix_internal use the moved state. */
int sdfibm_evolve(sdfibm_ctx* ctx, double time, double dt, const double* gravity);

/* Contacts between solids inside evolve (SolidCloud::solidSolidInteract and
 * solidSolidCollision, solidcloud.cpp:477-519; UGrid pair search,
 * libcollision/ugrid.cpp:5-77; geometry of collision.cpp:7-32).  Every finite
 * solid is its bounding sphere (IShape::m_radiusB of its shape), a plane solid
 * the half-space local y < 0.  Force on each partner along the line of centres
 * (plane: along the plane normal), no torque:
 *     f = max(0, stiffness * overlap - damping * v_n),
 * v_n = the normal relative velocity of the contact point (> 0 separating).
 * The reference's law is stiffness 1e4, damping 0; at its HEAD the search grid
 * is built with a bounding radius of -1 (solidcloud.cpp:74-75, 245) and finds no
 * pairs, which is stiffness = damping = 0 here (the default: no contacts).
 * Evaluated every sub-step from the state at its start; a solid's contacts are
 * summed in ascending partner index, then the planes.  Unlike the reference,
 * planes are not put in the search grid (it inserts a plane by its reference
 * point only) but tested against every finite solid, and shapes other than
 * sphere/circle use their bounding sphere instead of the reference's table
 * lookup (which maps unknown shape names to Plane, collision.h:11-15). */
int sdfibm_set_contact(sdfibm_ctx* ctx, double stiffness, double damping);

/* n_steps coupled steps of the reference's solver loop (src/main.cpp:66-84
 * without the PISO solve): interact -> U -= Fs dt -> evolve (with contacts if
 * set) -> fixInternal, each step a replay of one HIP graph captured at the
 * first call (re-captured when dt, gravity or any setting changes); no host
 * synchronisation between the steps.  Before the replays one interact of the
 * current state computes the plane solids' cell sets, which the replays keep:
 * plane solids must have motion mask b000000 and zero velocity and spin
 * (SDFIBM_ERR_STATE otherwise, and if a plane's pose changes during the
 * replays).  Capacity errors of any step are reported after the last one; the
 * step whose interact failed and every later step leave the solids and U as the
 * last valid step left them (its fields and force/torque are not valid), so the
 * caller can call run_steps again, which grows the buffers first.  ft: [ns*6] force/torque
 * of the last step's interact, or NULL.  time: unused (no time-dependent
 * motions or forcers).  sdfibm_timing's total is the replays' device time. */
int sdfibm_run_steps(sdfibm_ctx* ctx, double time, double dt, int64_t n_steps, const double* gravity, double* ft);

/* n_steps interacts of the current state (time, dt) as replays of one captured
 * HIP graph (captured at the first call, again when dt or any setting changes),
 * without host synchronisation between them: what a caller that keeps U and the
 * solids on the device issues per step.  One eager interact first computes the
 * plane solids' cell sets and the buffer sizes.  Errors of any replay are
 * reported after the last.  ft: [ns*6] force/torque of the last, or NULL.
 * sdfibm_timing's total is the replays' device time. */
int sdfibm_interact_steps(sdfibm_ctx* ctx, double time, double dt, int64_t n_steps, double* ft);

/* Device -> host copy of the solids' current records (after evolve). */
int sdfibm_get_solids(sdfibm_ctx* ctx, sdfibm_solid* out, int64_t n);

/* ---- learned volume-fraction surrogate (no counterpart in the reference) ----
 * An MLP 11 -> 32 -> 32 -> 1 (ReLU) over a cut cell's 8 corner signed distances
 * / h and the unit SDF gradient at the cell centre; BF16 weights and
 * activations, FP32 accumulation, MFMA.  Parameters: 1473 FP32 values, W1[32][11]
 * b1[32] W2[32][32] b2[32] w3[32] b3.  Labels are the reference method's own
 * fractions (geometrictools.cpp:47-72) on cells cut by random spheres and
 * ellipsoids. */
#define SDFIBM_SURROGATE_PARAMS 1473

typedef struct {
    uint64_t seed;
    int64_t n_train;   /* training samples generated on the device */
    int64_t n_val;     /* validation samples (another seed) */
    int64_t batch;
    int32_t epochs;
    int32_t reserved;
    double lr;         /* Adam learning rate, cosine-decayed to 0 */
} sdfibm_surrogate_train_config;

/* Train on the GPU (forward/backward kernels, Adam).  params: in = initial
 * values, out = trained values.  report (may be NULL): [0] last batch MSE,
 * [1] validation MSE and [2] max |error| through the deployed MFMA inference,
 * [3] seconds. */
int sdfibm_surrogate_train(const sdfibm_surrogate_train_config* cfg, float* params, double* report);

/* n samples of the training distribution: features [n*16] (8 corner distances
 * in cells, the centre gradient, zeros) and labels [n]. */
int sdfibm_surrogate_samples(uint64_t seed, int64_t n, float* features, float* labels);

/* The deployed MFMA inference on feature rows [n*16] -> out [n]. */
int sdfibm_surrogate_eval(const float* params, const float* features, int64_t n, float* out);

/* Gradient of 0.5 * sum (net(x) - y)^2 w.r.t. the parameters (the training
 * kernels' forward and backward pass; straight-through for the BF16 rounding).
 * sum_sq (may be NULL) receives sum (net(x) - y)^2. */
int sdfibm_surrogate_grad(const float* params, const float* features, const float* labels, int64_t n,
                          float* grad, double* sum_sq);

/* Use the surrogate in the next interact calls of a FP32 context on a
 * rectilinear 3-D mesh (params NULL: back to the exact fractions).  A cut cell
 * falls back to the exact fraction when the prediction lies in [band_lo,
 * band_hi] and the centre gradient is more than max_angle_deg away from the
 * corner-difference gradient (max_angle_deg < 0: whenever the prediction lies
 * in the band), and always for boundary cells, 2-D cases and shapes other than
 * spheres and ellipsoids. */
int sdfibm_set_surrogate(sdfibm_ctx* ctx, const float* params, double band_lo, double band_hi,
                         double max_angle_deg);

/* Cut cells the surrogate served in the last interact. */
int sdfibm_surrogate_count(sdfibm_ctx* ctx, int64_t* n);

/* Device -> host copy of U as AoS FP64 [N*3]. */
int sdfibm_get_velocity(sdfibm_ctx* ctx, double* U_aos);

/* Device -> host copies of the solid fields as FP64 (any pointer may be NULL):
 * As [N], Fs [N*3] AoS, Ts [N], Ct [N] (Ct stored as the reference's double field). */
int sdfibm_get_fields(sdfibm_ctx* ctx, double* As, double* Fs_aos, double* Ts, double* Ct);

/* Device pointers of the resident fields (element type per precision: float or
 * double; Ct is int32).  For zero-copy use by a device-resident caller. */
int sdfibm_field_pointers(sdfibm_ctx* ctx, void** U3 /* [3]: Ux,Uy,Uz */, void** As,
                          void** Fs3 /* [3] */, void** Ts, int32_t** Ct);

/* Candidate cells of one solid from the last interact, in the reference's
 * summation order: ALL_INSIDE ascending, CENTER_INSIDE ascending,
 * CENTER_OUTSIDE ascending (solidcloud.cpp:367-374).  counts[0..2] get the
 * three set sizes, counts[3] the seed cell (-1: none); cells/alpha (may be
 * NULL) get up to `capacity` entries.  For parity tests. */
int sdfibm_candidates(sdfibm_ctx* ctx, int64_t solid, int64_t* counts, int64_t* cells,
                      double* alpha, int64_t capacity);

/* Totals over all solids of the candidate sets of the last interact (also
 * after run_steps: its last step): out[0] = ALL_INSIDE pairs, out[1] = cut
 * pairs (CENTER_INSIDE + CENTER_OUTSIDE); their sum is the number of
 * (solid, cell) updates, sum_s |cellids_s| of solidcloud.cpp:403-422.  For
 * benchmarks. */
int sdfibm_pair_counts(sdfibm_ctx* ctx, int64_t* out);

/* Sizes of the context's step buffers (diagnostics and tests): out[0] = cells the
 * tile-code buffer holds, out[1] = solids the per-solid buffers hold.  interact grows
 * the code buffer when moving solids' tiles outgrow it.  Returns the number of values
 * written (<= n, at most 2).  No counterpart in the reference. */
int sdfibm_buffer_capacity(sdfibm_ctx* ctx, int64_t* out, int n);

/* Phase timing of the structured-mesh interact on (on != 0) or off (the
 * default).  On, every eager interact records HIP events between its phases for
 * sdfibm_last_timing; each record orders the kernels around it, which costs about
 * 25 us per step on a 256^3 / 10,000-solid scene, so it is off unless asked for.
 * sdfibm_autotune turns it on for its own timed steps.  No counterpart in the
 * reference. */
int sdfibm_set_phase_timing(sdfibm_ctx* ctx, int32_t on);

/* Timing of the phases of the last interact (milliseconds, HIP events on the
 * context's stream; structured meshes: only with sdfibm_set_phase_timing on,
 * zeros otherwise): [0] setup+binning, [1] per-solid tiles (classification,
 * cut cells and forcing, plane labelling, and the wait for the big-tile pass),
 * [2] field pass, [3] total, [4] tile classification alone, [5] cut cells +
 * forcing of the tiles of the on-chip tile pass, [6] the big-tile pass (on the
 * second stream, beside [4] and [5]; 0 without big tiles).  Returns the number of
 * values written (<= n, at most 7). */
int sdfibm_last_timing(sdfibm_ctx* ctx, double* ms, int n);

/* Run-time variants of the step (no counterpart in the reference): the cut-cell
 * pass of the tiles of the on-chip tile pass with one wave per solid
 * (cells_wave 1), 256 threads per solid (0) or by the mean tile size (-1, the
 * default); the big tiles' cut-cell pass in chunks of big_chunk candidates
 * (64..16384, default 1024) over big_grid workgroups per CU (1..32, default 4).
 * Results are the same up to the order of the FP64 force/torque sums. */
int sdfibm_set_tuning(sdfibm_ctx* ctx, int32_t cells_wave, int32_t big_chunk, int32_t big_grid);

/* Autotuner (no counterpart in the reference): times the variants above on the
 * context's solids and U (interact runs several times; the fields then hold the
 * current state's step) and keeps the fastest, searching one knob at a time --
 * per size bin: the small tiles' cut-cell pass, then, when big tiles can occur,
 * the big pass's chunk size and grid.  cache_path (may be NULL): a JSON-lines
 * file keyed by the mesh, precision and the histogram of the solids' tile sizes;
 * a matching line is applied without timing, a miss appends one.  chosen (may be
 * NULL) receives the choice as JSON text (cells_wave, big_chunk, big_grid, ms,
 * source "cache" or "timed", variants_timed). */
int sdfibm_autotune(sdfibm_ctx* ctx, double dt, const char* cache_path, char* chosen, int64_t chosen_cap);

/* Launch the work of interact on a caller-provided hipStream_t (NULL = the
 * context's own stream). */
int sdfibm_set_stream(sdfibm_ctx* ctx, void* hip_stream);

const char* sdfibm_last_error(const sdfibm_ctx* ctx);
const char* sdfibm_status_string(int status);
int sdfibm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SDFIBM_HIP_H */
