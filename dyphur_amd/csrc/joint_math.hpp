// joint_math.hpp — XPBD joint constraints on two bodies' register state
// (core/src/xpbd_solver.cpp:60-407 helpers and constraints, 523-556
// solve_joint), shared by the gfx950 solver kernels and the host build of
// the same code.  Compile with -ffp-contract=off: every float operation
// rounds on its own, in the reference's order.  A body with inv_mass == 0
// is read but never written.
#pragma once

#include "dy_math.hpp"

namespace dy {

enum : uint32_t { kJointFixed = 0u, kJointRevolute = 1u, kJointPrismatic = 2u, kJointBall = 3u };

// One joint's parameters (core/include/core/articulation.hpp:16-42).
struct JointIn {
    float apx, apy, apz;  // anchor in the parent's frame
    float acx, acy, acz;  // anchor in the child's frame
    float lx, ly, lz;     // axis in the parent's frame (revolute, prismatic)
    float lo, hi;         // limits; lo > hi disables them
    float target_pos, target_vel, stiffness, damping;
    float compliance_pos, compliance_ang;
    uint32_t type;
};

// atan(k / 8), k = 0..8, in float64.
DY_HD double atan_k8(int k) {
    switch (k) {
        case 0: return 0.0;
        case 1: return 0.12435499454676144;
        case 2: return 0.24497866312686414;
        case 3: return 0.35877067027057225;
        case 4: return 0.4636476090008061;
        case 5: return 0.5585993153435624;
        case 6: return 0.6435011087932844;
        case 7: return 0.7188299996216245;
        default: return 0.7853981633974483;
    }
}

// The correctly rounded float atan2: evaluated in float64 to about 2^-51
// relative and rounded once, so the result is the float nearest the exact
// angle except when that angle lies within 2^-51 of a midpoint between two
// floats.  The ratio t = min/max of |y|, |x| is reduced to u = (t - c) /
// (1 + t c) around the nearest c = k/8 (|u| <= 1/16), atan(u) by its odd
// Taylor series to u^17 (remainder below 2^-76 relative), then the octant
// (pi/2 - a), the half plane (pi - a) and the sign of y.  IEEE special
// cases as the C library's atan2.
DY_HD float atan2_cr(float yf, float xf) {
    const double kPi = 3.141592653589793, kPi2 = 1.5707963267948966;
    if (yf != yf || xf != xf) return yf + xf;
    const bool yneg = __builtin_signbit(yf) != 0, xneg = __builtin_signbit(xf) != 0;
    const double ay = __builtin_fabs(static_cast<double>(yf)), ax = __builtin_fabs(static_cast<double>(xf));
    const double inf = __builtin_inf();
    double a;
    if (ay == 0.0) {
        a = xneg ? kPi : 0.0;
    } else if (ax == 0.0) {
        a = kPi2;
    } else if (ay == inf || ax == inf) {
        if (ay == inf && ax == inf) a = xneg ? 0.75 * kPi : 0.25 * kPi;
        else if (ay == inf) a = kPi2;
        else a = xneg ? kPi : 0.0;
    } else {
        const bool swap = ay > ax;
        const double t = swap ? ax / ay : ay / ax;
        const int k = static_cast<int>(t * 8.0 + 0.5);
        const double c = static_cast<double>(k) * 0.125;
        const double u = (t - c) / (1.0 + t * c);
        const double u2 = u * u;
        double p = 1.0 / 17.0;
        p = -1.0 / 15.0 + u2 * p;
        p = 1.0 / 13.0 + u2 * p;
        p = -1.0 / 11.0 + u2 * p;
        p = 1.0 / 9.0 + u2 * p;
        p = -1.0 / 7.0 + u2 * p;
        p = 1.0 / 5.0 + u2 * p;
        p = -1.0 / 3.0 + u2 * p;
        a = atan_k8(k) + (u + u * (u2 * p));
        if (swap) a = kPi2 - a;
        if (xneg) a = kPi - a;
    }
    return static_cast<float>(yneg ? -a : a);
}

// Body-local v in world axes (qrot_to_world, xpbd_solver.cpp:62-73).
DY_HD void to_world(const BodyState &b, float vx, float vy, float vz, float &ox, float &oy, float &oz) {
    const M3 M = body_axes(b.qw, b.qx, b.qy, b.qz);
    ox = M.m[0][0] * vx + M.m[1][0] * vy + M.m[2][0] * vz;
    oy = M.m[0][1] * vx + M.m[1][1] * vy + M.m[2][1] * vz;
    oz = M.m[0][2] * vx + M.m[1][2] * vy + M.m[2][2] * vz;
}

// inv_mass + (r x n) . I_w^-1 (r x n)  (gen_inv_mass, :90-98).
DY_HD float gen_inv_mass(const BodyState &b, float rx, float ry, float rz, float nx, float ny, float nz) {
    const float cx = ry * nz - rz * ny, cy = rz * nx - rx * nz, cz = rx * ny - ry * nx;
    float kx, ky, kz;
    inv_inertia_world(b, cx, cy, cz, kx, ky, kz);
    return b.im + (cx * kx + cy * ky + cz * kz);
}

// Positional impulse sign * dl * n at lever arm r, with the velocity change
// dx / h (apply_pos_impulse, :103-133).
DY_HD void pos_impulse(BodyState &b, float sign, float dl, float nx, float ny, float nz, float rx, float ry,
                       float rz, float inv_dt) {
    const float im = b.im;
    if (im == 0.f) return;
    const float dx = sign * im * dl * nx, dy = sign * im * dl * ny, dz = sign * im * dl * nz;
    b.px += dx; b.py += dy; b.pz += dz;
    b.vx += dx * inv_dt; b.vy += dy * inv_dt; b.vz += dz * inv_dt;
    const float cx = ry * nz - rz * ny, cy = rz * nx - rx * nz, cz = rx * ny - ry * nx;
    float kx, ky, kz;
    inv_inertia_world(b, cx, cy, cz, kx, ky, kz);
    const float ax = sign * dl * kx, ay = sign * dl * ky, az = sign * dl * kz;
    rotate_by(b, ax, ay, az);
    b.wx += ax * inv_dt; b.wy += ay * inv_dt; b.wz += az * inv_dt;
}

// Anchor coincidence (solve_pos_constraint, :137-164).
DY_HD void joint_pos(BodyState &A, BodyState &B, float rax, float ray, float raz, float rbx, float rby, float rbz,
                     float alpha, float inv_dt) {
    const float eps = 1e-7f;
    const float cx = (A.px + rax) - (B.px + rbx);
    const float cy = (A.py + ray) - (B.py + rby);
    const float cz = (A.pz + raz) - (B.pz + rbz);
    const float c2 = cx * cx + cy * cy + cz * cz;
    if (c2 < eps * eps) return;
    const float cl = sqrtf(c2);
    const float nx = cx / cl, ny = cy / cl, nz = cz / cl;
    const float wa = gen_inv_mass(A, rax, ray, raz, nx, ny, nz);
    const float wb = gen_inv_mass(B, rbx, rby, rbz, nx, ny, nz);
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -cl / w;
    pos_impulse(A, +1.f, dl, nx, ny, nz, rax, ray, raz, inv_dt);
    pos_impulse(B, -1.f, dl, nx, ny, nz, rbx, rby, rbz, inv_dt);
}

// Anchor coincidence across the world axis u only (solve_transverse_pos,
// :169-196).
DY_HD void joint_transverse(BodyState &A, BodyState &B, float rax, float ray, float raz, float rbx, float rby,
                            float rbz, float ux, float uy, float uz, float alpha, float inv_dt) {
    const float cx = (A.px + rax) - (B.px + rbx);
    const float cy = (A.py + ray) - (B.py + rby);
    const float cz = (A.pz + raz) - (B.pz + rbz);
    const float along = cx * ux + cy * uy + cz * uz;
    const float tx = cx - along * ux, ty = cy - along * uy, tz = cz - along * uz;
    const float t2 = tx * tx + ty * ty + tz * tz;
    if (t2 < 1e-14f) return;
    const float tl = sqrtf(t2);
    const float nx = tx / tl, ny = ty / tl, nz = tz / tl;
    const float wa = gen_inv_mass(A, rax, ray, raz, nx, ny, nz);
    const float wb = gen_inv_mass(B, rbx, rby, rbz, nx, ny, nz);
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -tl / w;
    pos_impulse(A, +1.f, dl, nx, ny, nz, rax, ray, raz, inv_dt);
    pos_impulse(B, -1.f, dl, nx, ny, nz, rbx, rby, rbz, inv_dt);
}

// Rotation correction along world vector c, |c| = error angle; angular
// velocities change by the rotation times inv_dt (solve_ang_constraint,
// :200-230).
DY_HD void joint_ang(BodyState &A, BodyState &B, float cx, float cy, float cz, float alpha, float inv_dt) {
    const float c2 = cx * cx + cy * cy + cz * cz;
    if (c2 < 1e-14f) return;
    const float cl = sqrtf(c2);
    const float nx = cx / cl, ny = cy / cl, nz = cz / cl;
    float ax, ay, az, bx, by, bz;
    inv_inertia_world(A, nx, ny, nz, ax, ay, az);
    inv_inertia_world(B, nx, ny, nz, bx, by, bz);
    const float wa = nx * ax + ny * ay + nz * az;
    const float wb = nx * bx + ny * by + nz * bz;
    const float w = wa + wb + alpha;
    if (w < 1e-10f) return;
    const float dl = -cl / w;
    rotate_by(A, -dl * ax, -dl * ay, -dl * az);
    rotate_by(B, dl * bx, dl * by, dl * bz);
    if (A.im != 0.f) {
        A.wx += (-dl * ax) * inv_dt; A.wy += (-dl * ay) * inv_dt; A.wz += (-dl * az) * inv_dt;
    }
    if (B.im != 0.f) {
        B.wx += (dl * bx) * inv_dt; B.wy += (dl * by) * inv_dt; B.wz += (dl * bz) * inv_dt;
    }
}

// Relative orientation locked (solve_fixed_ang, :234-252).
DY_HD void joint_fixed_ang(BodyState &A, BodyState &B, float alpha, float inv_dt) {
    const float aw = A.qw, ax = A.qx, ay = A.qy, az = A.qz;
    const float bw = B.qw, bx = B.qx, by = B.qy, bz = B.qz;
    const float rw = aw * bw + ax * bx + ay * by + az * bz;
    float rx = -ax * bw + aw * bx - az * by + ay * bz;
    float ry = -ay * bw + az * bx + aw * by - ax * bz;
    float rz = -az * bw - ay * bx + ax * by + aw * bz;
    if (rw < 0.f) { rx = -rx; ry = -ry; rz = -rz; }
    float cx, cy, cz;
    to_world(A, rx, ry, rz, cx, cy, cz);
    cx *= 2.f; cy *= 2.f; cz *= 2.f;
    joint_ang(A, B, cx, cy, cz, alpha, inv_dt);
}

// Hinge axes aligned (solve_revolute_ang, :257-273).
DY_HD void joint_hinge(BodyState &A, BodyState &B, float lx, float ly, float lz, float alpha, float inv_dt) {
    float ax, ay, az, bx, by, bz;
    to_world(A, lx, ly, lz, ax, ay, az);
    to_world(B, lx, ly, lz, bx, by, bz);
    if (ax * bx + ay * by + az * bz < 0.f) { bx = -bx; by = -by; bz = -bz; }
    const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    joint_ang(A, B, cx, cy, cz, alpha, inv_dt);
}

// Joint coordinate: offset along the axis (prismatic) or hinge angle
// (revolute) (joint_current_value, :278-318).
DY_HD float joint_value(const BodyState &A, const BodyState &B, float lx, float ly, float lz, uint32_t type) {
    float ux, uy, uz;
    to_world(A, lx, ly, lz, ux, uy, uz);
    if (type == kJointPrismatic) return (B.px - A.px) * ux + (B.py - A.py) * uy + (B.pz - A.pz) * uz;
    float px, py, pz;
    if (fabsf(lx) < fabsf(ly) && fabsf(lx) < fabsf(lz)) {
        const float len = sqrtf(ly * ly + lz * lz);
        px = 0.f; py = -lz / len; pz = ly / len;
    } else if (fabsf(ly) < fabsf(lz)) {
        const float len = sqrtf(lx * lx + lz * lz);
        px = lz / len; py = 0.f; pz = -lx / len;
    } else {
        const float len = sqrtf(lx * lx + ly * ly);
        px = -ly / len; py = lx / len; pz = 0.f;
    }
    float rx, ry, rz, cx, cy, cz;
    to_world(A, px, py, pz, rx, ry, rz);
    to_world(B, px, py, pz, cx, cy, cz);
    const float sx = ry * cz - rz * cy, sy = rz * cx - rx * cz, sz = rx * cy - ry * cx;
    const float sn = sx * ux + sy * uy + sz * uz;
    const float cs = rx * cx + ry * cy + rz * cz;
    return atan2_cr(sn, cs);
}

// One-sided limit on the joint coordinate (apply_limit, :322-352).
DY_HD void joint_limit(BodyState &A, BodyState &B, const JointIn &j, float inv_dt) {
    if (j.lo > j.hi) return;
    const float v = joint_value(A, B, j.lx, j.ly, j.lz, j.type);
    float viol = 0.f;
    if (v < j.lo) viol = v - j.lo;
    else if (v > j.hi) viol = v - j.hi;
    if (fabsf(viol) < 1e-7f) return;
    float ux, uy, uz;
    to_world(A, j.lx, j.ly, j.lz, ux, uy, uz);
    if (j.type == kJointPrismatic) {
        const float wa = gen_inv_mass(A, 0.f, 0.f, 0.f, ux, uy, uz);
        const float wb = gen_inv_mass(B, 0.f, 0.f, 0.f, ux, uy, uz);
        const float w = wa + wb;
        if (w < 1e-10f) return;
        const float dl = -viol / w;
        pos_impulse(A, +1.f, dl, ux, uy, uz, 0.f, 0.f, 0.f, inv_dt);
        pos_impulse(B, -1.f, dl, ux, uy, uz, 0.f, 0.f, 0.f, inv_dt);
    } else {
        joint_ang(A, B, viol * ux, viol * uy, viol * uz, 0.f, 0.f);
    }
}

// PD motor of a revolute or prismatic joint: one velocity impulse along the
// axis per solve (apply_motor, :356-407); other types have none.
DY_HD void joint_motor(BodyState &A, BodyState &B, const JointIn &j, float h) {
    if (j.type != kJointRevolute && j.type != kJointPrismatic) return;
    if (j.stiffness < 1e-6f && j.damping < 1e-6f) return;
    float ux, uy, uz;
    to_world(A, j.lx, j.ly, j.lz, ux, uy, uz);
    float vrel;
    if (j.type == kJointPrismatic) {
        const float vx = B.vx - A.vx, vy = B.vy - A.vy, vz = B.vz - A.vz;
        vrel = vx * ux + vy * uy + vz * uz;
    } else {
        const float ox = B.wx - A.wx, oy = B.wy - A.wy, oz = B.wz - A.wz;
        vrel = ox * ux + oy * uy + oz * uz;
    }
    float pos_err = 0.f;
    if (j.stiffness > 1e-6f) pos_err = j.target_pos - joint_value(A, B, j.lx, j.ly, j.lz, j.type);
    const float vel_err = j.target_vel - vrel;
    const float ima = A.im, imb = B.im;
    if (j.type == kJointPrismatic) {
        const float w = ima + imb;
        if (w < 1e-10f) return;
        const float imp = (j.stiffness * pos_err + j.damping * vel_err) * h / w;
        if (ima != 0.f) { A.vx -= ima * imp * ux; A.vy -= ima * imp * uy; A.vz -= ima * imp * uz; }
        if (imb != 0.f) { B.vx += imb * imp * ux; B.vy += imb * imp * uy; B.vz += imb * imp * uz; }
    } else {
        float ax, ay, az, bx, by, bz;
        inv_inertia_world(A, ux, uy, uz, ax, ay, az);
        inv_inertia_world(B, ux, uy, uz, bx, by, bz);
        const float wa = ux * ax + uy * ay + uz * az;
        const float wb = ux * bx + uy * by + uz * bz;
        const float w = wa + wb;
        if (w < 1e-10f) return;
        const float imp = (j.stiffness * pos_err + j.damping * vel_err) * h / w;
        if (ima != 0.f) { A.wx -= imp * ax; A.wy -= imp * ay; A.wz -= imp * az; }
        if (imb != 0.f) { B.wx += imp * bx; B.wy += imp * by; B.wz += imp * bz; }
    }
}

// One joint constraint, A = parent, B = child (solve_joint, :524-556).  The
// lever arms are taken once, before this joint's corrections.
DY_HD void solve_joint(const JointIn &j, BodyState &A, BodyState &B, float dt) {
    float rax, ray, raz, rbx, rby, rbz;
    to_world(A, j.apx, j.apy, j.apz, rax, ray, raz);
    to_world(B, j.acx, j.acy, j.acz, rbx, rby, rbz);
    const float alpha_pos = j.compliance_pos / (dt * dt);
    const float alpha_ang = j.compliance_ang / (dt * dt);
    const float inv_dt = 1.f / dt;
    if (j.type == kJointPrismatic) {
        float ux, uy, uz;
        to_world(A, j.lx, j.ly, j.lz, ux, uy, uz);
        joint_fixed_ang(A, B, alpha_ang, inv_dt);
        joint_transverse(A, B, rax, ray, raz, rbx, rby, rbz, ux, uy, uz, alpha_pos, inv_dt);
        joint_limit(A, B, j, inv_dt);
    } else {
        joint_pos(A, B, rax, ray, raz, rbx, rby, rbz, alpha_pos, inv_dt);
        if (j.type == kJointFixed) {
            joint_fixed_ang(A, B, alpha_ang, inv_dt);
        } else if (j.type == kJointRevolute) {
            joint_hinge(A, B, j.lx, j.ly, j.lz, alpha_ang, inv_dt);
            joint_limit(A, B, j, inv_dt);
        }
    }
}

}  // namespace dy
