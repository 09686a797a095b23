/*
 * sdfibm_oracle.h -- CPU FP64 oracle for the solid-fluid coupling step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch CPU restatement of the
 * reference's coupling step (SolidCloud::interact and everything under it).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker (or the timed CPU baseline); the product
 * library never links it.
 *
 * It works on a general OpenFOAM polyMesh (points, faces, owner, neighbour)
 * and derives cell->face, cell->point, cell->cell addressing and the face /
 * cell geometry the way OpenFOAM's primitiveMesh does.
 */
#ifndef SDFIBM_ORACLE_H
#define SDFIBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shape ids; the same numbering as the product's sdfibm_hip.h. */
enum {
    ORC_SPHERE = 0,
    ORC_ELLIPSOID = 1,
    ORC_CIRCLE = 2,
    ORC_ELLIPSE = 3,
    ORC_BOX = 4,
    ORC_RECTANGLE = 5,
    ORC_PLANE = 6,
    ORC_CIRCLE_TAIL = 7,
    ORC_CIRCLE_TWOTAIL = 8,
    ORC_CSG = 9,
    ORC_NSHAPES = 10
};

/* CSG shapes (the product's sdfibm_csg_shape, same layout): up to 8 primitives
 * (box, ellipsoid, finite cylinder, sphere) in the solid's local frame combined
 * left to right with the reference's operators (sdf.h:133-142).  A CSG solid's
 * prm[0] is the index of its shape in the table given to orc_set_csg. */
enum { ORC_CSG_BOX = 0, ORC_CSG_ELLIPSOID = 1, ORC_CSG_CYLINDER = 2, ORC_CSG_SPHERE = 3 };
enum { ORC_CSG_UNION = 0, ORC_CSG_DIFFERENCE = 1, ORC_CSG_INTERSECTION = 2 };
typedef struct {
    int32_t kind, op, axis, reserved;
    double center[3];
    double size[3];
} orc_csg_prim;
typedef struct {
    int32_t n_prims, reserved;
    double volume;
    double moi[3];
    double radius_b;
    orc_csg_prim prim[8];
} orc_csg_shape;
/* The CSG table the oracle's CSG solids refer to (copied; process-wide). */
void orc_set_csg(const orc_csg_shape* shapes, int64_t n);

/* One rigid solid: shape parameters as written in a solidDict plus state.
 * prm meaning per shape (unused slots ignored):
 *   SPHERE, CIRCLE      : prm[0]=radius, prm[1..3]=com
 *   ELLIPSOID           : prm[0..2]=radiusa,b,c           (no com: ellipsoid.h ignores it)
 *   ELLIPSE, RECTANGLE  : prm[0..1]=radiusa,b, prm[3..5]=com
 *   BOX                 : prm[0..2]=radiusa,b,c, prm[3..5]=com
 *   PLANE               : none
 *   CIRCLE_TAIL/TWOTAIL : prm[0]=radius, prm[1]=ratio, prm[2]=thickness, prm[3..5]=com
 * Layout is identical to sdfibm_solid in include/sdfibm_hip.h (160 bytes). */
typedef struct {
    int32_t shape;
    int32_t flags;
    double prm[6];
    double center[3];
    double q[4];      /* w, x, y, z */
    double vel[3];
    double omega[3];
} orc_solid;

typedef struct orc_mesh orc_mesh;

/* Build a mesh from polyMesh arrays (copied).  Returns NULL on bad input. */
orc_mesh* orc_mesh_create(int64_t n_points, const double* points,
                          int64_t n_faces, const int64_t* face_offsets, const int32_t* face_points,
                          const int32_t* owner, int64_t n_internal, const int32_t* neighbour);
/* Same for a single-block hex mesh given only its points (i-fastest), faces
 * generated in blockMesh's order (internal by owner, then patches xmin..zmax). */
orc_mesh* orc_mesh_create_structured(int64_t nx, int64_t ny, int64_t nz, const double* points);
void orc_mesh_destroy(orc_mesh* m);
int64_t orc_mesh_ncells(const orc_mesh* m);
/* Export derived geometry: cc[3N], V[N], fc[3F], Sf[3F] (any may be NULL). */
void orc_mesh_geometry(const orc_mesh* m, double* cc, double* V, double* fc, double* Sf);
/* Export cell->point addressing (ascending labels): offsets[N+1], labels. */
int64_t orc_mesh_cell_points(const orc_mesh* m, int64_t* offsets, int32_t* labels);

/* Shape kernels (local coordinates), exposed for closed-form tests. */
int orc_is_inside(const orc_solid* s, const double* p_world);
double orc_signed_distance(const orc_solid* s, const double* p_world);
void orc_world_to_local(const orc_solid* s, const double* p_world, double* p_local);
/* Fraction of cell `cell` inside solid s (calcCellVolume / V). */
double orc_cell_fraction(const orc_mesh* m, const orc_solid* s, int64_t cell, int two_d);
/* Nearest cell centre (ties: lowest cell index). */
int64_t orc_find_nearest_cell(const orc_mesh* m, const double* p);

/* Per-solid candidate capture (all optional).
 *   counts[4*s+0..3] = #ALL_INSIDE, #CENTER_INSIDE, #CENTER_OUTSIDE, seed cell (-1: none)
 *   cells / alpha: concatenated lists in the reference's summation order
 *   (ALL_INSIDE ascending, CENTER_INSIDE ascending, CENTER_OUTSIDE ascending). */
typedef struct {
    int64_t* counts;
    int64_t* cells;
    double* alpha;
    int64_t capacity;
    int64_t used;
} orc_capture;

/* One interact() call.
 *   U [3N] in, As/Ts/Ct [N] and Fs [3N] out (overwritten), force_torque [6*ns] out.
 *   mode 0: fast bookkeeping (stamped visited array, shared vertex cache);
 *   mode 1: reference-cost bookkeeping (N-sized visited array and std::set per solid,
 *           per-solid hash-map vertex cache).  Both give identical results.
 * Returns 0, or -1 on bad arguments, -2 if the capture buffer is too small. */
int orc_interact(const orc_mesh* m, int two_d, double rho_f, double dt,
                 int64_t ns, const orc_solid* solids, const double* U,
                 double* As, double* Fs, double* Ts, double* Ct,
                 double* force_torque, int mode, orc_capture* cap);

/* Solid-only partial interact for timing subsets: processes solids[s0:s1]
 * into already-zeroed fields (no zero-fill, no clip). */
int orc_interact_range(const orc_mesh* m, int two_d, double rho_f, double dt,
                       int64_t ns, const orc_solid* solids, int64_t s0, int64_t s1,
                       const double* U, double* As, double* Fs, double* Ts, double* Ct,
                       double* force_torque, int mode);

/* The steps around interact: U -= Fs*dt (main.cpp:70) and fixInternal
 * (solidcloud.cpp:288-301).  U, Fs AoS [3N]; Ct as the double field. */
void orc_update_velocity(int64_t n, double* U, const double* Fs, double dt);
void orc_fix_internal(const orc_mesh* m, double* U, const double* Ct, int64_t ns, const orc_solid* solids);

/* Rigid-body state that SolidCloud::evolve carries per solid beyond orc_solid
 * (solid.h:28-55): density, Motion01Mask bits (bit 0..5 = vx vy vz wx wy wz
 * free; -1 = no motion constraint), the first-fluid-step flag and the previous
 * step's fluid force and torque. */
typedef struct {
    double rho;
    int32_t mask;
    int32_t first_fluid;
    double ff_old[3];
    double ft_old[3];
} orc_dyn;

/* Shape volume and principal moments of inertia per unit density (the
 * m_volume / m_moi of each shape constructor; plane: 0 and identity). */
void orc_mass_properties(const orc_solid* s, double* volume, double* moi3);

/* SolidCloud::evolve(time, dt) (solidcloud.cpp:521-562): n_sub sub-steps of
 * clear -> AB2 fluid force (solid.h:164-174) -> buoyant gravity
 * (solidcloud.cpp:466-475) -> Solid::move (solid.h:182-202), then the fluid
 * force is stored as old.  fluid_ft: [ns*6] from the last interact (x rho_f).
 * Forcers and collisions are not part of this restatement (the reference's
 * collision grid is built with a negative bounding radius, solidcloud.cpp:74-75). */
void orc_evolve(int64_t ns, orc_solid* solids, orc_dyn* dyn, const double* fluid_ft, double time, double dt,
                const double* gravity, double rho_f);

/* Contacts (this library's law; the reference's geometry): IShape::m_radiusB of
 * a shape (plane: 0); the contact force on every solid by brute force over all
 * pairs -- f = max(0, k overlap - cdamp v_n) along the line of centres for two
 * finite solids (bounding spheres), along the plane normal for a plane; summed
 * over finite partners in ascending index, then the planes; out [3*ns]; and
 * evolve with those forces added after gravity in every sub-step.  k = cdamp =
 * 0 gives orc_evolve. */
double orc_bounding_radius(const orc_solid* s);
void orc_contact_forces(int64_t ns, const orc_solid* solids, double k, double cdamp, double* out);
void orc_evolve_contact(int64_t ns, orc_solid* solids, orc_dyn* dyn, const double* fluid_ft, double time, double dt,
                        const double* gravity, double rho_f, double k, double cdamp);

#ifdef __cplusplus
}
#endif
#endif
