// mc_physics.hpp — per-particle arithmetic shared by the HIP kernels and the
// host setup code (host callers: source-cell search and the cell search of
// the loader).  Every expression keeps the reference's operation order and the
// build uses -ffp-contract=off, so a history computed here on the GPU is
// bit-identical to the CPU restatement of the same algorithm.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/mc_api.h"

#if defined(__HIPCC__)
#define MC_HD __host__ __device__ inline
#else
#define MC_HD inline
#endif

namespace mc {

// Constants.h:11-13
constexpr double kEpsFloat = 1.1920928955078125e-07;     // FLT_EPSILON
constexpr double kMaxFloat = 3.4028234663852886e+38;     // FLT_MAX
constexpr double kMaxFloatLess = 0.9 * kMaxFloat;

// Random.cpp:92-101 (MCNP5 63-bit generator #5)
constexpr uint64_t kRnMult = 3512401965023503517ULL;
constexpr uint64_t kRnStride = 152917ULL;
constexpr uint64_t kRnMask = ~0ULL >> 1;
constexpr double kRnNorm = 1.0 / 9223372036854775808.0;  // 2^-63

// Urand, Random.cpp:121-126, on a per-particle state
MC_HD double urand(uint64_t& s)
{
    s = (kRnMult * s) & kRnMask;
    return (double)s * kRnNorm;
}

MC_HD double bits_to_double(uint64_t b)
{
    union { uint64_t u; double d; } v;
    v.u = b;
    return v.d;
}
MC_HD uint64_t double_to_bits(double d)
{
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}

// ln x for positive normal x: x = 2^e m, m in (1/sqrt2, sqrt2], ln m =
// 2 atanh(s), s = (m-1)/(m+1), odd series to s^22, e ln2 in Cody-Waite parts.
MC_HD double mc_log(double x)
{
    uint64_t b = double_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double r = 1.0 / 23.0;
    r = 1.0 / 21.0 + z * r;
    r = 1.0 / 19.0 + z * r;
    r = 1.0 / 17.0 + z * r;
    r = 1.0 / 15.0 + z * r;
    r = 1.0 / 13.0 + z * r;
    r = 1.0 / 11.0 + z * r;
    r = 1.0 / 9.0 + z * r;
    r = 1.0 / 7.0 + z * r;
    r = 1.0 / 5.0 + z * r;
    r = 1.0 / 3.0 + z * r;
    const double lnm = 2.0 * s + (2.0 * s) * (z * r);
    const double de = (double)e;
    return de * 0.6931471804855391 + (lnm + de * 7.440617110012397e-11);
}

// cos(2 pi xi), sin(2 pi xi): exact quadrant reduction on xi, then Taylor
// series of sin (to theta^17) and cos (to theta^18) on theta <= pi/4.
MC_HD void sincos_2pi(double xi, double& c, double& s)
{
    const double t = 4.0 * xi;
    int q = (int)t;
    const double r = t - (double)q;
    q &= 3;
    const bool swap = r > 0.5;
    const double th = (swap ? (1.0 - r) : r) * 1.5707963267948966;
    const double z = th * th;
    double ps = -8.22063524662433e-18;
    ps = 2.8114572543455206e-15 + z * ps;
    ps = -7.647163731819816e-13 + z * ps;
    ps = 1.6059043836821613e-10 + z * ps;
    ps = -2.505210838544172e-08 + z * ps;
    ps = 2.7557319223985893e-06 + z * ps;
    ps = -0.0001984126984126984 + z * ps;
    ps = 0.008333333333333333 + z * ps;
    ps = -0.16666666666666666 + z * ps;
    const double sn = th + th * (z * ps);
    double pc = -1.5619206968586225e-16;
    pc = 4.779477332387385e-14 + z * pc;
    pc = -1.1470745597729725e-11 + z * pc;
    pc = 2.08767569878681e-09 + z * pc;
    pc = -2.755731922398589e-07 + z * pc;
    pc = 2.48015873015873e-05 + z * pc;
    pc = -0.001388888888888889 + z * pc;
    pc = 0.041666666666666664 + z * pc;
    pc = -0.5 + z * pc;
    const double cs = 1.0 + z * pc;
    const double c0 = swap ? sn : cs;
    const double s0 = swap ? cs : sn;
    if (q == 0) { c = c0; s = s0; }
    else if (q == 1) { c = -s0; s = c0; }
    else if (q == 2) { c = -c0; s = -s0; }
    else { c = s0; s = -c0; }
}

// Algorithm.cpp:46-64: largest i with v[i] < x (-1 below, n-1 above)
MC_HD int binary_search(double x, const double* v, int n)
{
    int left = 0, right = n - 1;
    while (left <= right) {
        const int mid = (left + right) / 2;
        if (v[mid] < x) left = mid + 1;
        else right = mid - 1;
    }
    return right;
}

// Algorithm.cpp:103-105
MC_HD double interpolate(double x, double x1, double x2, double y1, double y2)
{
    return (x - x2) / (x1 - x2) * y1 + (x - x1) / (x2 - x1) * y2;
}

// Algorithm.cpp:16-38
MC_HD double geometry_quad(double a, double b, double c)
{
    const double D = b * b - 4.0 * a * c;
    if (D <= 0.0) return kMaxFloat;
    const double sqrtD = sqrt(D);
    const double ai = 0.5 / a;
    double r1 = ai * (-1.0 * b - sqrtD);
    double r2 = ai * (-1.0 * b + sqrtD);
    if (r1 < 0) r1 = kMaxFloat;
    if (r2 < 0) r2 = kMaxFloat;
    return fmin(r1, r2);
}

struct Vec3 {
    double x, y, z;
};

// Surface::eval, Geometry.cpp:29-69
MC_HD double surface_eval(int type, const double* c, Vec3 p)
{
    switch (type) {
    case MC_SURF_PLANE_X: return p.x - c[0];
    case MC_SURF_PLANE_Y: return p.y - c[0];
    case MC_SURF_PLANE_Z: return p.z - c[0];
    case MC_SURF_PLANE: return c[0] * p.x + c[1] * p.y + c[2] * p.z - c[3];
    case MC_SURF_SPHERE: {
        const double xt = p.x - c[0], yt = p.y - c[1], zt = p.z - c[2];
        return xt * xt + yt * yt + zt * zt - c[4];
    }
    case MC_SURF_CYL_X: {
        const double yt = p.y - c[0], zt = p.z - c[1];
        return yt * yt + zt * zt - c[4];
    }
    case MC_SURF_CYL_Z: {
        const double xt = p.x - c[0], yt = p.y - c[1];
        return xt * xt + yt * yt - c[4];
    }
    default: return 0.0;
    }
}

MC_HD double plane_axis_distance(double pos, double dir, double loc)
{
    // Geometry.cpp:76-126
    if (fabs(dir) > kEpsFloat) {
        const double dist = (loc - pos) / dir;
        if (dist > 0.0) return dist;
        return kMaxFloat;
    }
    return kMaxFloat;
}

// Surface::distance, Geometry.cpp:76-188
MC_HD double surface_distance(int type, const double* c, Vec3 p, Vec3 u)
{
    switch (type) {
    case MC_SURF_PLANE_X: return plane_axis_distance(p.x, u.x, c[0]);
    case MC_SURF_PLANE_Y: return plane_axis_distance(p.y, u.y, c[0]);
    case MC_SURF_PLANE_Z: return plane_axis_distance(p.z, u.z, c[0]);
    case MC_SURF_PLANE: {
        const double denom = c[0] * u.x + c[1] * u.y + c[2] * u.z;
        if (fabs(denom) > kEpsFloat) {
            const double dist = (c[3] - c[0] * p.x - c[1] * p.y - c[2] * p.z) / denom;
            if (dist > 0.0) return dist;
            return kMaxFloat;
        }
        return kMaxFloat;
    }
    case MC_SURF_SPHERE: {
        const double b = 2.0 * ((p.x - c[0]) * u.x + (p.y - c[1]) * u.y + (p.z - c[2]) * u.z);
        return geometry_quad(1.0, b, surface_eval(type, c, p));
    }
    case MC_SURF_CYL_X: {
        const double a = 1.0 - u.x * u.x;
        const double b = 2.0 * ((p.y - c[0]) * u.y + (p.z - c[1]) * u.z);
        return geometry_quad(a, b, surface_eval(type, c, p));
    }
    case MC_SURF_CYL_Z: {
        const double a = 1.0 - u.z * u.z;
        const double b = 2.0 * ((p.x - c[0]) * u.x + (p.y - c[1]) * u.y);
        return geometry_quad(a, b, surface_eval(type, c, p));
    }
    default: return kMaxFloat;
    }
}

// Surface::reflect, Geometry.cpp:195-222 (quadrics do not reflect)
MC_HD Vec3 surface_reflect(int type, const double* c, Vec3 u)
{
    switch (type) {
    case MC_SURF_PLANE_X: u.x = -u.x; break;
    case MC_SURF_PLANE_Y: u.y = -u.y; break;
    case MC_SURF_PLANE_Z: u.z = -u.z; break;
    case MC_SURF_PLANE: {
        const double K = (c[0] * u.x + c[1] * u.y + c[2] * u.z);
        const Vec3 q = {u.x - c[4] * K, u.y - c[5] * K, u.z - c[6] * K};
        return q;
    }
    default: break;
    }
    return u;
}

// scatter_direction, Algorithm.cpp:67-101, given the azimuth's cos and sin
MC_HD Vec3 scatter_direction(Vec3 d, double mu0, double cos_azi, double sin_azi)
{
    const double Ac = sqrt(1.0 - mu0 * mu0);
    Vec3 o;
    if (d.z != 1.0) {
        const double B = sqrt(1.0 - d.z * d.z);
        const double C = Ac / B;
        o.x = d.x * mu0 + (d.x * d.z * cos_azi - d.y * sin_azi) * C;
        o.y = d.y * mu0 + (d.y * d.z * cos_azi + d.x * sin_azi) * C;
        o.z = d.z * mu0 - cos_azi * Ac * B;
    } else {
        const double B = sqrt(1.0 - d.y * d.y);
        const double C = Ac / B;
        o.x = d.x * mu0 + (d.x * d.y * cos_azi - d.z * sin_azi) * C;
        o.z = d.z * mu0 + (d.z * d.y * cos_azi + d.x * sin_azi) * C;
        o.y = d.y * mu0 - cos_azi * Ac * B;
    }
    return o;
}

// ReactionScatter::sample, Reaction.cpp:27-118: elastic scatter of a
// particle of energy E (eV) and direction u off a free gas of mass ratio A at
// 293.6 K (constant cross section).  mu0 isotropic in the centre of mass, the
// target's speed and cosine from the reference's two-branch rejection (its
// Eq. 37-41), centre-of-mass kinematics, then the lab speed and energy with
// Particle.cpp:42-56's constants.  The reference leaves the order of the two
// draws in -log(U) - log(U) cos^2 unspecified (C++ operand order); they are
// taken left to right.  cos(pi/2 xi) is cos(2 pi (xi / 4)).
MC_HD void free_gas_scatter(double A, double& E, Vec3& u, uint64_t& rng)
{
    const double mu0 = 2.0 * urand(rng) - 1.0;
    const double beta = sqrt(2.0659834e-11 * A);
    const double v = 13831.5926439 * sqrt(E) * 100.0;
    const double y = beta * v;
    double V, mut;
    do {
        double x;
        if (urand(rng) < 2.0 / (2.0 + 1.7724538509055159 * y)) {
            const double u1 = urand(rng);
            const double u2 = urand(rng);
            x = sqrt(-mc_log(u1 * u2));
        } else {
            double c, sn;
            sincos_2pi(0.25 * urand(rng), c, sn);
            const double u1 = urand(rng);
            const double u2 = urand(rng);
            x = sqrt(-mc_log(u1) - mc_log(u2) * c * c);
        }
        V = x / beta;
        mut = 2.0 * urand(rng) - 1.0;
    } while (urand(rng) > sqrt(v * v + V * V - 2.0 * v * V * mut) / (v + V));
    double ca, sa;
    sincos_2pi(urand(rng), ca, sa);
    const Vec3 nd = scatter_direction(u, mut, ca, sa);       // the target's direction
    const Vec3 vl = {v * u.x, v * u.y, v * u.z};
    const Vec3 cm = {(vl.x + A * (nd.x * V)) / (1.0 + A), (vl.y + A * (nd.y * V)) / (1.0 + A),
                     (vl.z + A * (nd.z * V)) / (1.0 + A)};  // centre-of-mass velocity
    Vec3 vc = {vl.x - cm.x, vl.y - cm.y, vl.z - cm.z};
    const double sc = sqrt(vc.x * vc.x + vc.y * vc.y + vc.z * vc.z);
    const Vec3 dc = {vc.x / sc, vc.y / sc, vc.z / sc};
    sincos_2pi(urand(rng), ca, sa);
    const Vec3 dn = scatter_direction(dc, mu0, ca, sa);
    const Vec3 w = {sc * dn.x + cm.x, sc * dn.y + cm.y, sc * dn.z + cm.z};
    const double sp = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    E = 5.2270376e-13 * sp * sp;
    u = {w.x / sp, w.y / sp, w.z / sp};
}

// Per-history stream: RN_init_particle (Random.cpp:196-204) starts history
// `key` at seed * g^(key * stride).  jump[256 j + v] = g^(stride v 256^j)
// mod 2^63, so the start is one multiply per byte of key (eight table reads
// that do not depend on each other) instead of RN_skip_ahead's
// square-and-multiply; products mod 2^63 do not depend on their order.
MC_HD uint64_t history_seed(uint64_t seed, uint64_t key, const uint64_t* jump)
{
    uint64_t f[8];
    for (int j = 0; j < 8; ++j) f[j] = jump[256 * j + ((key >> (8 * j)) & 255u)];
    uint64_t g = seed & kRnMask;
    for (int j = 0; j < 8; ++j) g = (g * f[j]) & kRnMask;
    return g;
}

}  // namespace mc
