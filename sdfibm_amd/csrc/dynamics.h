// dynamics.h -- rigid-body update of the solids (SolidCloud::evolve,
// src/solidcloud.cpp:521-562; Solid::move, src/solid.h:182-202), FP64, with
// the reference's operation order: OpenFOAM vector/tensor/quaternion algebra
// written out (quaternion product, quaternion::R(), tensor & tensor).
#pragma once

#include "shapes.h"

namespace sdfibm_dev {

// Per-solid rigid-body state on the device beyond the solid record.
struct SolidDyn {
    double rho;         // solid density (IMaterial::getRho)
    int mask;           // Motion01Mask bits 0..5 = vx vy vz wx wy wz free; -1 = unconstrained
    int first_fluid;    // Solid::first_fluid_step
    double ff_old[3];   // fluid force and torque of the previous step (storeOldForce)
    double ft_old[3];
    double vol, moi[3]; // CSG solids: the mass properties their shape record declares
};

// Shape volume and principal moments of inertia per unit density: the m_volume
// and diagonal m_moi each shape constructor sets (plane: 0 and the identity).
__host__ __device__ inline void mass_properties(const sdfibm_solid& s, double& vol, double moi[3]) {
    const double PI = 3.14159265358979323846;
    vol = 0.0;
    moi[0] = moi[1] = moi[2] = 1.0;
    switch (s.shape) {
    case SDFIBM_SHAPE_SPHERE: {  // sphere.h:21-25
        const double r = s.prm[0], r2 = r * r;
        vol = 4.0 / 3.0 * PI * r2 * r;
        moi[0] = moi[1] = moi[2] = 0.4 * vol * r2;
        break;
    }
    case SDFIBM_SHAPE_ELLIPSOID: {  // ellipsoid.h:25-30
        const double a = s.prm[0], b = s.prm[1], c = s.prm[2];
        vol = 4.0 / 3.0 * PI * a * b * c;
        moi[0] = 0.2 * vol * (b * b + c * c);
        moi[1] = 0.2 * vol * (a * a + c * c);
        moi[2] = 0.2 * vol * (a * a + b * b);
        break;
    }
    case SDFIBM_SHAPE_CIRCLE:
    case SDFIBM_SHAPE_CIRCLE_TAIL:
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: {  // circle.h:22-25; the tails do not count
        const double r = s.prm[0], r2 = r * r;
        vol = PI * r2;
        moi[0] = moi[1] = moi[2] = 0.5 * vol * r2;
        break;
    }
    case SDFIBM_SHAPE_ELLIPSE: {  // ellipse.h:24-28
        const double a = s.prm[0], b = s.prm[1];
        vol = PI * a * b;
        moi[0] = moi[1] = moi[2] = 0.25 * vol * (a * a + b * b);
        break;
    }
    case SDFIBM_SHAPE_BOX: {  // box.h:22-27
        const double a = s.prm[0], b = s.prm[1], c = s.prm[2];
        vol = 8.0 * a * b * c;
        moi[0] = 1.0 / 3.0 * vol * (b * b + c * c);
        moi[1] = 1.0 / 3.0 * vol * (a * a + c * c);
        moi[2] = 1.0 / 3.0 * vol * (b * b + a * a);
        break;
    }
    case SDFIBM_SHAPE_RECTANGLE: {  // rectangle.h:25-29
        const double a = s.prm[0], b = s.prm[1];
        vol = 4.0 * a * b;
        moi[0] = moi[1] = moi[2] = 1.0 / 3.0 * vol * (a * a + b * b);
        break;
    }
    default: break;
    }
}

// Mass, inverse mass and principal inverse moments of a solid (Solid::setMaterial,
// solid.h:94-103: mass = V rho, mass_inv = (1/V)/rho, moi_inv = inv(moi)/rho).
struct BodyConst {
    double mass, mass_inv, Mi[3];
};

__host__ __device__ inline BodyConst body_const(const sdfibm_solid& s, const SolidDyn& d) {
    double vol, moi[3];
    mass_properties(s, vol, moi);
    if (s.shape == SDFIBM_SHAPE_CSG) {
        vol = d.vol;
        for (int a = 0; a < 3; ++a) moi[a] = d.moi[a];
    }
    const double volinv = vol > 0 ? 1.0 / vol : 0.0;
    BodyConst b;
    b.mass = vol * d.rho;
    b.mass_inv = volinv / d.rho;
    for (int a = 0; a < 3; ++a) b.Mi[a] = (1.0 / moi[a]) / d.rho;
    return b;
}

// The loads of one sub-step before contacts: clearForceAndTorque, the AB2 mix of
// the fluid force and torque (addMidFluidForceAndTorque, solid.h:164-174; the
// first call takes old = current) and the buoyant gravity (addMidEnvironment,
// solidcloud.cpp:466-475).
__host__ __device__ inline void substep_loads(SolidDyn& d, const d3& ff, const d3& ft, const d3& gv, double rho_f,
                                              const BodyConst& b, d3& force, d3& torque) {
    if (d.first_fluid) {
        d.ff_old[0] = ff.x, d.ff_old[1] = ff.y, d.ff_old[2] = ff.z;
        d.ft_old[0] = ft.x, d.ft_old[1] = ft.y, d.ft_old[2] = ft.z;
        d.first_fluid = 0;
    }
    const d3 ffo{d.ff_old[0], d.ff_old[1], d.ff_old[2]}, fto{d.ft_old[0], d.ft_old[1], d.ft_old[2]};
    force = d3{0, 0, 0} + (1.5 * ff - 0.5 * ffo);
    torque = d3{0, 0, 0} + (1.5 * ft - 0.5 * fto);
    const d3 gp = ((d.rho - rho_f) / d.rho) * gv;
    force = force + b.mass * gp;
}

// Solid::move (solid.h:181-202) over one sub-step of length dts.
__host__ __device__ inline void move_solid(sdfibm_solid& s, const SolidDyn& d, const BodyConst& b, const d3& force,
                                           const d3& torque, double dts) {
    d3 v{s.vel[0], s.vel[1], s.vel[2]}, w{s.omega[0], s.omega[1], s.omega[2]};
    const d3 v_old = v, w_old = w;
    v = v + dts * (b.mass_inv * force);
    const double qw = s.q[0];
    const d3 qv{s.q[1], s.q[2], s.q[3]};
    // quaternion::R()
    const double w2 = qw * qw, x2 = qv.x * qv.x, y2 = qv.y * qv.y, z2 = qv.z * qv.z;
    const double txy = 2 * qv.x * qv.y, twz = 2 * qw * qv.z, txz = 2 * qv.x * qv.z;
    const double twy = 2 * qw * qv.y, tyz = 2 * qv.y * qv.z, twx = 2 * qw * qv.x;
    const double R[9] = {w2 + x2 - y2 - z2, txy - twz, txz + twy, txy + twz, w2 - x2 + y2 - z2,
                         tyz - twx,         txz - twy, tyz + twx, w2 - x2 - y2 + z2};
    // (R & diag(Mi)) & R^T, the tensor products written out in OpenFOAM's summation order
    double T1[9], Mw[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            T1[3 * i + j] = R[3 * i] * (j == 0 ? b.Mi[0] : 0.0) + R[3 * i + 1] * (j == 1 ? b.Mi[1] : 0.0) +
                            R[3 * i + 2] * (j == 2 ? b.Mi[2] : 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Mw[3 * i + j] = T1[3 * i] * R[3 * j] + T1[3 * i + 1] * R[3 * j + 1] + T1[3 * i + 2] * R[3 * j + 2];
    const d3 aw{Mw[0] * torque.x + Mw[1] * torque.y + Mw[2] * torque.z,
                Mw[3] * torque.x + Mw[4] * torque.y + Mw[5] * torque.z,
                Mw[6] * torque.x + Mw[7] * torque.y + Mw[8] * torque.z};
    w = w + dts * aw;
    if (d.mask >= 0) {  // Motion01Mask::constraint (motion01mask.h:60-64)
        v = {v.x * ((d.mask >> 0) & 1), v.y * ((d.mask >> 1) & 1), v.z * ((d.mask >> 2) & 1)};
        w = {w.x * ((d.mask >> 3) & 1), w.y * ((d.mask >> 4) & 1), w.z * ((d.mask >> 5) & 1)};
    }
    const d3 c = d3{s.center[0], s.center[1], s.center[2]} + dts * (0.5 * (v + v_old));
    // q += (0.5 quaternion(0.5 (w + w_old)) * q) dt, then normalise
    const d3 hv = 0.5 * (0.5 * (w + w_old));
    const double hw = 0.5 * 0.0;
    const double dw = hw * qw - dot(hv, qv);
    const d3 dv = (hw * qv + qw * hv) + cross(hv, qv);
    double nw = qw + dw * dts;
    d3 nv = qv + dts * dv;
    const double qm = sqrt(nw * nw + (nv.x * nv.x + nv.y * nv.y + nv.z * nv.z));
    nw = nw / qm;
    nv = {nv.x / qm, nv.y / qm, nv.z / qm};
    s.vel[0] = v.x, s.vel[1] = v.y, s.vel[2] = v.z;
    s.omega[0] = w.x, s.omega[1] = w.y, s.omega[2] = w.z;
    s.center[0] = c.x, s.center[1] = c.y, s.center[2] = c.z;
    s.q[0] = nw, s.q[1] = nv.x, s.q[2] = nv.y, s.q[3] = nv.z;
}

// Solid::storeOldForce (solid.h:143-149).
__host__ __device__ inline void store_old_force(SolidDyn& d, const d3& ff, const d3& ft) {
    d.ff_old[0] = ff.x, d.ff_old[1] = ff.y, d.ff_old[2] = ff.z;
    d.ft_old[0] = ft.x, d.ft_old[1] = ft.y, d.ft_old[2] = ft.z;
}

// One SolidCloud::evolve of one solid without contacts: n_sub sub-steps of AB2
// fluid force + buoyant gravity + Solid::move, then the fluid force is stored
// as old.  fft: this step's fluid force and torque (x rho_f); g: gravity.
__host__ __device__ inline void evolve_solid(sdfibm_solid& s, SolidDyn& d, const double* fft, int n_sub, double dt,
                                             const double* g, double rho_f) {
    const double dts = dt / n_sub;
    const BodyConst b = body_const(s, d);
    const d3 ff{fft[0], fft[1], fft[2]}, ft{fft[3], fft[4], fft[5]};
    const d3 gv{g[0], g[1], g[2]};
    for (int it = 0; it < n_sub; ++it) {
        d3 force, torque;
        substep_loads(d, ff, ft, gv, rho_f, b, force, torque);
        move_solid(s, d, b, force, torque, dts);
    }
    store_old_force(d, ff, ft);
}

// ---------------------------------------------------------------------------
// Contacts between solids (the collision stage of SolidCloud::evolve,
// solidcloud.cpp:477-519, collision.cpp:7-32).  Geometry: the reference's --
// every finite solid is its bounding sphere of radius IShape::m_radiusB, a
// plane is the half-space local y < 0 with normal = the local +y axis.  Force
// law (this library's; the reference hard-codes 1e4 * overlap * normal, which
// is k = 1e4, c = 0 here): a linear spring-dashpot along the line of centres,
//   f = max(0, k delta - c v_n) n,
// delta the overlap, v_n the normal relative velocity of the contact point
// (positive when separating), no tangential force and hence no torque.

// IShape::m_radiusB of the shapes (plane and unknown: 0).
__host__ __device__ inline double bounding_radius(const sdfibm_solid& s) {
    switch (s.shape) {
    case SDFIBM_SHAPE_SPHERE:  // sphere.h:28
    case SDFIBM_SHAPE_CIRCLE:  // circle.h:27
        return s.prm[0];
    case SDFIBM_SHAPE_ELLIPSOID:  // ellipsoid.h:32
    case SDFIBM_SHAPE_BOX:        // box.h:29
        return fmax(fmax(s.prm[0], s.prm[1]), s.prm[2]);
    case SDFIBM_SHAPE_ELLIPSE:    // ellipse.h:30
    case SDFIBM_SHAPE_RECTANGLE:  // rectangle.h:31
        return fmax(s.prm[0], s.prm[1]);
    case SDFIBM_SHAPE_CIRCLE_TAIL:     // circle_tail.h:28,33: 2 (ratio + 1) r / 2
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL:  // circle_twotail.h:26,31
        return 2 * ((s.prm[1] + 1) * 0.5 * s.prm[0]);
    default:
        return 0.0;
    }
}

struct ContactLaw {
    double k, c;  // stiffness, damping
};

// Overlap of finite solids a and b (bounding radii ra, rb); > 0 means contact.
__host__ __device__ inline double pair_overlap(const sdfibm_solid& a, double ra, const sdfibm_solid& b, double rb) {
    const d3 ab{b.center[0] - a.center[0], b.center[1] - a.center[1], b.center[2] - a.center[2]};
    return ra + rb - sqrt(dot(ab, ab));
}

// Contact force on a from b (sphereSphereCollision geometry: normal from a to
// b, contact point 0.5 (ca + cb + (ra - rb) n)).  Exactly antisymmetric: the
// call with a and b swapped returns the negated vector bit for bit.
__host__ __device__ inline d3 pair_force(const sdfibm_solid& a, double ra, const sdfibm_solid& b, double rb,
                                         ContactLaw law) {
    const d3 ca{a.center[0], a.center[1], a.center[2]}, cb{b.center[0], b.center[1], b.center[2]};
    const d3 ab = cb - ca;
    const double dist = sqrt(dot(ab, ab));
    const double delta = ra + rb - dist;
    const double den = dist + 1e-300;  // Foam::normalised: v / (mag(v) + VSMALL)
    const d3 n{ab.x / den, ab.y / den, ab.z / den};
    const d3 cp = 0.5 * ((ca + cb) + (ra - rb) * n);
    const d3 va = d3{a.vel[0], a.vel[1], a.vel[2]} + cross(d3{a.omega[0], a.omega[1], a.omega[2]}, cp - ca);
    const d3 vb = d3{b.vel[0], b.vel[1], b.vel[2]} + cross(d3{b.omega[0], b.omega[1], b.omega[2]}, cp - cb);
    const double vn = dot(vb - va, n);
    const double fm = law.k * delta - law.c * vn;
    if (!(delta > 0) || !(fm > 0)) return d3{0, 0, 0};
    return (-fm) * n;
}

// Contact force on finite solid s (bounding radius r) from plane p
// (planeSphereCollision geometry: overlap r - (local y of the centre), contact
// point centre - r n).
__host__ __device__ inline d3 plane_force(const sdfibm_solid& s, double r, const sdfibm_solid& p, ContactLaw law) {
    const double qw = p.q[0], qx = p.q[1], qy = p.q[2], qz = p.q[3];
    // the plane's local +y axis in world axes: column 1 of quaternion::R()
    const d3 n{2 * qx * qy - 2 * qw * qz, qw * qw - qx * qx + qy * qy - qz * qz, 2 * qy * qz + 2 * qw * qx};
    const d3 cs{s.center[0], s.center[1], s.center[2]}, cpl{p.center[0], p.center[1], p.center[2]};
    const double p2s = dot(n, cs - cpl);
    const double delta = r - p2s;
    const d3 cp = cs - r * n;
    const d3 vs = d3{s.vel[0], s.vel[1], s.vel[2]} + cross(d3{s.omega[0], s.omega[1], s.omega[2]}, cp - cs);
    const d3 vp = d3{p.vel[0], p.vel[1], p.vel[2]} + cross(d3{p.omega[0], p.omega[1], p.omega[2]}, cp - cpl);
    const double vn = dot(vs - vp, n);
    const double fm = law.k * delta - law.c * vn;
    if (!(delta > 0) || !(fm > 0)) return d3{0, 0, 0};
    return fm * n;
}

}  // namespace sdfibm_dev
