// device_math.hpp -- per-pair and per-agent arithmetic shared by the gfx950 kernels.
//
// Every function restates one reference routine; the FP64 instantiations keep the
// reference's operation order (the translation unit is compiled with
// -ffp-contract=off), the float instantiations are the mixed-precision pair path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "params.hpp"

namespace crm {

template <typename T>
__device__ __forceinline__ T dsqrt(T x);
template <>
__device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }

template <typename T>
__device__ __forceinline__ T dexp(T x);
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float dexp<float>(float x) { return expf(x); }

template <typename T>
__device__ __forceinline__ T dpow(T x, T y);
template <>
__device__ __forceinline__ double dpow<double>(double x, double y) { return pow(x, y); }
template <>
__device__ __forceinline__ float dpow<float>(float x, float y) { return powf(x, y); }

template <typename T>
__device__ __forceinline__ T dmin(T a, T b);  // Rust f{32,64}::min: NaN-ignoring
template <>
__device__ __forceinline__ double dmin<double>(double a, double b) { return fmin(a, b); }
template <>
__device__ __forceinline__ float dmin<float>(float a, float b) { return fminf(a, b); }
template <typename T>
__device__ __forceinline__ T dmax(T a, T b);
template <>
__device__ __forceinline__ double dmax<double>(double a, double b) { return fmax(a, b); }
template <>
__device__ __forceinline__ float dmax<float>(float a, float b) { return fmaxf(a, b); }

// x^e for a small non-negative integer e, by binary powering.
template <typename T>
__device__ __forceinline__ T ipow(T x, int e) {
    T r = T(1);
    T b = x;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

template <typename T>
__device__ __forceinline__ T powe(T x, T e, int ie) {
    return ie >= 0 ? ipow<T>(x, ie) : dpow<T>(x, e);
}

// ---------------------------------------------------------------------------
// Point potentials.  z = own - ext.  Writes the force on `own`; the force on
// `ext` is its exact negation for every potential below (IEEE negation is exact
// and the reference returns (dir * F, -dir * F)).  Returns true on CalcError.
//
// `cut` overrides the cutoff test: -1 decides from z, 0 forces "beyond the
// cutoff", 1 forces "inside".  The float pair path uses it when the FP64
// decision was taken from the double positions (the force jumps at the cutoff,
// so the decision has to be the reference's).
// ---------------------------------------------------------------------------

// Float pair path: distance and its reciprocal from one v_rsq_f32, exp from
// v_exp_f32 (relative errors ~1e-7, far inside the mixed-precision budget).
struct FastDist {
    float d2, dist, inv;
};
template <int D>
__device__ __forceinline__ FastDist fast_dist(const float* z) {
    float d2 = 0.0f;
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    const float inv = __builtin_amdgcn_rsqf(d2);
    return FastDist{d2, d2 * inv, inv};
}

template <int D>
__device__ __forceinline__ void morse_fast(const float* z, float two_radius, float cutoff, float strength,
                                           float stiffness, int cut, float* f) {
    const FastDist q = fast_dist<D>(z);
    const bool beyond = cut < 0 ? (q.dist > cutoff) : (cut == 0);
    if (beyond || q.d2 == 0.0f) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = 0.0f;
        return;
    }
    const float e = __expf((-stiffness) * (q.dist - two_radius));
    const float force = ((((-2.0f) * strength) * stiffness) * e) * (1.0f - e);
    const float fi = force * q.inv;
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = z[k] * fi;
}

// calculate_morse_interaction, interaction.rs:151-182
template <typename T, int D>
__device__ __forceinline__ void morse(const T* z, T two_radius, T cutoff, T strength, T stiffness, int cut, T* f) {
    if constexpr (std::is_same<T, float>::value) {
        morse_fast<D>(z, two_radius, cutoff, strength, stiffness, cut, f);
        return;
    }
    T d2 = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    T dist = dsqrt<T>(d2);
    const bool beyond = cut < 0 ? (dist > cutoff) : (cut == 0);
    if (beyond || dist == T(0)) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return;
    }
    T e = dexp<T>((-stiffness) * (dist - two_radius));
    T force = ((((T(-2)) * strength) * stiffness) * e) * (T(1) - e);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = (z[k] / dist) * force;
}

// Float Mie pair path: 1/r from v_rsq_f32 and one Newton step serves sigma / r and
// the direction, en / sigma and em / sigma come precomputed (the FP64 path below
// keeps the reference's operations).  sigma / r enters to the power en + 1, so
// 1/r gets the Newton step: on the 4M-agent Mie lattice the 500-step RMS position
// drift against FP64 was 1.5e-5 with the bare v_rsq_f32, 1.05e-5 with the step and
// 8.0e-6 with sqrtf and divisions (a drift set by the few agents near unstable
// equilibria, which any change of rounding tips differently; the force error on
// the end state is 1.1e-6 RMS either way).
template <int D>
__device__ __forceinline__ bool mie_fast(const float* z, const PotParams<float>& P, int cut, float* f) {
    FastDist q;
    q.d2 = 0.0f;
#pragma unroll
    for (int k = 0; k < D; k++) q.d2 = q.d2 + z[k] * z[k];
    {
        const float y = __builtin_amdgcn_rsqf(q.d2);
        const float t = q.d2 * y;
        q.inv = fmaf(0.5f * y, fmaf(-t, y, 1.0f), y);  // y (3 - d2 y^2) / 2
        q.dist = q.d2 * q.inv;
    }
    const bool beyond = cut < 0 ? (q.dist > P.cutoff) : (cut == 0);
    if (q.d2 == 0.0f || beyond) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = 0.0f;
        return q.d2 == 0.0f;
    }
    const float sr = P.mie_sigma * q.inv;
    const float part =
        P.mie_en_s * powe<float>(sr, P.en + 1.0f, P.ien1) - P.mie_em_s * powe<float>(sr, P.em + 1.0f, P.iem1);
    const float force = fminf(P.mie_constant * P.strength * part, P.bound);
    const float fi = force * q.inv;
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = z[k] * fi;
    return false;
}

// MiePotential::calculate_force_between, interaction.rs:345-376.
// sigma and the Mie constant are pair-independent and precomputed on the host
// with the reference's expressions.
template <typename T, int D>
__device__ __forceinline__ bool mie(const T* z, const PotParams<T>& P, int cut, T* f) {
    if constexpr (std::is_same<T, float>::value) return mie_fast<D>(z, P, cut, f);
    T d2 = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    T r = dsqrt<T>(d2);
    if (r == T(0)) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return true;
    }
    const bool beyond = cut < 0 ? (r > P.cutoff) : (cut == 0);
    if (beyond) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return false;
    }
    T sr = P.mie_sigma / r;
    T part = P.en / P.mie_sigma * powe<T>(sr, P.en + T(1), P.ien1) -
             P.em / P.mie_sigma * powe<T>(sr, P.em + T(1), P.iem1);
    T force = dmin<T>(P.mie_constant * P.strength * part, P.bound);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = (z[k] / r) * force;
    return false;
}

// Float path: 1/r from v_rsq_f32 and one Newton step for sigma / r, 4 epsilon / r,
// bound / r and the direction (in place of sqrtf and six divisions).
template <int D>
__device__ __forceinline__ void bound_lj_fast(const float* z, const PotParams<float>& P, int cut, float* f) {
    float d2 = 0.0f;
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    const float y0 = __builtin_amdgcn_rsqf(d2);
    const float y = fmaf(0.5f * y0, fmaf(-(d2 * y0), y0, 1.0f), y0);  // 1 / r
    const float r = d2 * y;
    const float sr = P.sigma * y;
    const float val = (4.0f * P.epsilon * y) * (12.0f * ipow<float>(sr, 11) - 6.0f * ipow<float>(sr, 5));
    const float mx = P.bound * y;
    const bool inside = cut < 0 ? (P.cutoff >= r) : (cut == 1);
    const float q = inside ? 1.0f : 0.0f;
    const float m = fminf(mx, val);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = ((-(z[k] * y)) * q) * m;
}

// BoundLennardJones::calculate_force_between, interaction.rs:122-138
template <typename T, int D>
__device__ __forceinline__ void bound_lj(const T* z, const PotParams<T>& P, int cut, T* f) {
    if constexpr (std::is_same<T, float>::value) {
        bound_lj_fast<D>(z, P, cut, f);
        return;
    }
    T d2 = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    T r = dsqrt<T>(d2);
    T sr = P.sigma / r;
    T val = T(4) * P.epsilon / r * (T(12) * ipow<T>(sr, 11) - T(6) * ipow<T>(sr, 5));
    T mx = P.bound / r;
    const bool inside = cut < 0 ? (P.cutoff >= r) : (cut == 1);
    T q = inside ? T(1) : T(0);
    T m = dmin<T>(mx, val);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = ((-(z[k] / r)) * q) * m;
}

// CellSpecificInteraction, benchmarks/src/cell_sorting.rs:31-94 (sign +1) and
// examples/cell_sorting/src/main.rs:32-103 (sign -1).  `same` = equal species.
// Its discontinuity is the signum at relative_interaction_range * (R + R).
// Float path: 1/|z| from v_rsq_f32 and one Newton step gives the direction and
// (outside the core) 1 / sigma = 2R / r, in place of sqrtf and five divisions.
template <int D>
__device__ __forceinline__ void cell_sorting_fast(const float* z, bool same, const PotParams<float>& P, int cut,
                                                  float* f) {
    const float mrd = 0.3162277660168379f;
    float d2 = 0.0f;
    bool zero = true;
#pragma unroll
    for (int k = 0; k < D; k++) {
        d2 = d2 + z[k] * z[k];
        zero = zero && (z[k] == 0.0f);
    }
    if (zero) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = 0.0f;
        return;
    }
    const float y0 = __builtin_amdgcn_rsqf(d2);
    const float y = fmaf(0.5f * y0, fmaf(-(d2 * y0), y0, 1.0f), y0);  // 1 / |z|
    const float nz = d2 * y;
    const float core = P.radius * mrd;
    const bool in_core = nz < core;
    const float r = in_core ? core : nz;
    const float inv = P.two_radius * (in_core ? __builtin_amdgcn_rcpf(core) : y);  // 1 / sigma
    const float bound = 4.0f + inv;
    float spatial_cutoff;
    if (cut < 0) {
        const float sgn_arg = P.rel_range * P.two_radius - r;
        const float signum = (sgn_arg != sgn_arg) ? sgn_arg : copysignf(1.0f, sgn_arg);
        spatial_cutoff = (1.0f + signum) * 0.5f;
    } else {
        spatial_cutoff = cut ? 1.0f : 0.0f;
    }
    const float inv2 = inv * inv;
    const float strength = P.strength * fmaxf(fminf(inv2 - inv2 * inv2, bound), -bound);
    const float smax = fmaxf(strength, 0.0f), smin = fminf(strength, 0.0f);
#pragma unroll
    for (int k = 0; k < D; k++) {
        const float dir = z[k] * y;
        const float att = (dir * smax) * spatial_cutoff;
        const float rep = (dir * smin) * spatial_cutoff;
        const float o = same ? rep + att : rep;
        f[k] = P.sign < 0.0f ? -o : o;
    }
}

template <typename T, int D>
__device__ __forceinline__ void cell_sorting(const T* z, bool same, const PotParams<T>& P, int cut, T* f) {
    if constexpr (std::is_same<T, float>::value) {
        cell_sorting_fast<D>(z, same, P, cut, f);
        return;
    }
    const T mrd = T(0.3162277660168379);
    T d2 = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
    T nz = dsqrt<T>(d2);
    T r;
    if (!(nz < P.radius * mrd)) {
        r = nz;
    } else {
        bool zero = true;
#pragma unroll
        for (int k = 0; k < D; k++) zero = zero && (z[k] == T(0));
        if (zero) {
#pragma unroll
            for (int k = 0; k < D; k++) f[k] = T(0);
            return;
        }
        r = P.radius * mrd;
    }
    T sigma = r / P.two_radius;
    T bound = T(4) + T(1) / sigma;
    T spatial_cutoff;
    if (cut < 0) {
        T sgn_arg = P.rel_range * P.two_radius - r;
        T signum = (sgn_arg != sgn_arg) ? sgn_arg : copysign(T(1), sgn_arg);
        spatial_cutoff = (T(1) + signum) * T(0.5);
    } else {
        spatial_cutoff = cut ? T(1) : T(0);
    }
    T inv = T(1) / sigma;
    T inv2 = inv * inv;
    T strength = P.strength * dmax<T>(dmin<T>(inv2 - inv2 * inv2, bound), -bound);
    T smax = dmax<T>(strength, T(0)), smin = dmin<T>(strength, T(0));
#pragma unroll
    for (int k = 0; k < D; k++) {
        T dir = z[k] / nz;
        T att = (dir * smax) * spatial_cutoff;
        T rep = (dir * smin) * spatial_cutoff;
        T o = same ? rep + att : rep;
        f[k] = P.sign < T(0) ? -o : o;
    }
}

// CellSpecificInteraction of examples/cell_sorting/src/main-langevin.rs:44-92.
// R = R_own + R_ext, r = |z|, dir = z / r (zero at r == 0):
//   r <= R:       force = (repulsion * (R - r)) * dir
//   r <= cutoff:  force = ((((-a) * (r - R)) * (cutoff - r)) / cutoff) * exp(-1 * (r / R - 1)^2) * dir
//   otherwise     0
// with a = attraction of the (own, ext) species pair; it returns (-force, force),
// so the force on own is -force.  Continuous at R and at the cutoff.  The float
// path takes r and 1/r from one v_rsq_f32 and the exponential from v_exp_f32.
template <typename T, int D>
__device__ __forceinline__ void sorting_langevin(const T* z, T attraction, const PotParams<T>& P, int cut, T* f) {
    T d2 = T(0), r, inv;
    if constexpr (std::is_same<T, float>::value) {
        const FastDist q = fast_dist<D>(z);
        d2 = q.d2;
        r = q.dist;
        inv = q.inv;
    } else {
#pragma unroll
        for (int k = 0; k < D; k++) d2 = d2 + z[k] * z[k];
        r = dsqrt<T>(d2);
        inv = T(0);
    }
    if (d2 == T(0)) {  // dir = 0
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return;
    }
    const T R = P.two_radius;
    T s;
    if (r <= R) {
        s = P.strength * (R - r);
    } else if (cut < 0 ? (r <= P.cutoff) : (cut == 1)) {
        const T sm1 = r / R - T(1);
        T e;
        if constexpr (std::is_same<T, float>::value) e = __expf(-(sm1 * sm1));
        else e = dexp<T>(T(-1) * (sm1 * sm1));
        s = ((((-attraction) * (r - R)) * (P.cutoff - r)) / P.cutoff) * e;
    } else {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return;
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        if constexpr (std::is_same<T, float>::value) f[k] = -(s * (z[k] * inv));
        else f[k] = -(s * (z[k] / r));
    }
}

// Radius at which the potential is discontinuous (the cutoff test).
template <int POT, typename T>
__device__ __forceinline__ T cut_radius(const PotParams<T>& P) {
    if constexpr (POT == CRM_POTENTIAL_CELL_SORTING) return P.rel_range * P.two_radius;
    else return P.cutoff;
}

// Reference cutoff decision from FP64 distance r: 1 inside, 0 beyond.
template <int POT>
__device__ __forceinline__ int cut_decision(const PotParams<double>& P, double r) {
    if constexpr (POT == CRM_POTENTIAL_CELL_SORTING) {
        double a = P.rel_range * P.two_radius - r;
        return (a != a) ? 0 : (copysign(1.0, a) > 0.0 ? 1 : 0);
    } else if constexpr (POT == CRM_POTENTIAL_BOUND_LENNARD_JONES) {
        return (P.cutoff >= r) ? 1 : 0;
    } else {
        return (r > P.cutoff) ? 0 : 1;
    }
}

// One Interaction::calculate_force_between evaluation.  `strength` is the
// pair's interaction strength (Morse: the potential's strength, or that times
// the species-matrix entry; the Langevin sorting interaction: the species pair's
// attraction; computed by the caller), `same` tells the
// cell-sorting interaction whether both agents are of one species.
template <int POT, typename T, int D>
__device__ __forceinline__ bool point_force(const T* z, bool same, T strength, const PotParams<T>& P, int cut,
                                            T* f) {
    if constexpr (POT == CRM_POTENTIAL_MORSE || POT == CRM_POTENTIAL_MORSE_SPECIES) {
        morse<T, D>(z, P.two_radius, P.cutoff, strength, P.stiffness, cut, f);
        return false;
    } else if constexpr (POT == CRM_POTENTIAL_MIE) {
        return mie<T, D>(z, P, cut, f);
    } else if constexpr (POT == CRM_POTENTIAL_BOUND_LENNARD_JONES) {
        bound_lj<T, D>(z, P, cut, f);
        return false;
    } else if constexpr (POT == CRM_POTENTIAL_NONE) {
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = T(0);
        return false;
    } else if constexpr (POT == CRM_POTENTIAL_CELL_SORTING_LANGEVIN) {
        sorting_langevin<T, D>(z, strength, P, cut, f);
        return false;
    } else {
        cell_sorting<T, D>(z, same, P, cut, f);
        return false;
    }
}

// Strength of the (own, ext) species pair from a row of the matrix held in registers.
template <typename T>
__device__ __forceinline__ T pick4(const T* row, int k) {
    return k == 0 ? row[0] : (k == 1 ? row[1] : (k == 2 ? row[2] : row[3]));
}

// nearest_point_from_point_to_line, interaction.rs:497-509 (t and the point x).
template <typename T, int D>
__device__ __forceinline__ void nearest_point(const T* p, const T* a, const T* b, T* x, T& t) {
    T ab[D], ap[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        ab[k] = b[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    T num = T(0), den = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) {
        num = num + ab[k] * ap[k];
        den = den + ab[k] * ab[k];
    }
    T tt = num / den;
    if (tt < T(0)) tt = T(0);
    if (tt > T(1)) tt = T(1);
    t = tt;
#pragma unroll
    for (int k = 0; k < D; k++) x[k] = a[k] * (T(1) - tt) + b[k] * tt;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller normals.
// Counter {agent_lo, agent_hi, step_lo, (step_hi << 16) | block}, key = seed.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Writes the two normals of block b for (seed, agent, step).
__device__ __forceinline__ void gaussian_pair(uint64_t seed, uint64_t agent, uint64_t step, uint32_t b,
                                              double& z0, double& z1) {
    uint32_t c[4] = {(uint32_t)agent, (uint32_t)(agent >> 32), (uint32_t)step,
                     (uint32_t)(((step >> 32) << 16) | (uint64_t)b)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)c[1] << 32) | c[0];
    uint64_t d = ((uint64_t)c[3] << 32) | c[2];
    const double inv53 = 1.0 / 9007199254740992.0;
    double u1 = (double)((a >> 11) + 1) * inv53;
    double u2 = (double)(d >> 11) * inv53;
    double rad = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    double sn, cs;
    sincos(th, &sn, &cs);  // one argument reduction for both
    z0 = rad * cs;
    z1 = rad * sn;
}

// ---------------------------------------------------------------------------
// Grid helpers (cartesian_cuboid_n.rs:469-491, 594-631).
// ---------------------------------------------------------------------------

// num-traits to_usize on (p - min) / dx: ok iff -1 < x (NaN fails); trunc.
__host__ __device__ __forceinline__ bool axis_index(double p, double mn, double dx, uint32_t n, uint32_t& out) {
    double x = (p - mn) / dx;
    if (!(x > -1.0)) return false;
    if (!(x < 4294967296.0)) return false;
    double t = trunc(x);
    uint32_t i = t <= 0.0 ? 0u : (uint32_t)t;
    if (i >= n) return false;  // no such voxel: the cell stays where it is (update_mechanics.rs:430-453)
    out = i;
    return true;
}

// axis_index with the quotient estimated by a multiplication: (p - mn) * inv_dx
// is within a few units in the last place of the correctly rounded (p - mn) / dx
// (3e-16 relative, below 2e-6 absolute for indices under 2^32), so where its
// fractional part is further than 1e-4 from an integer both truncate to the same
// index; near integers, at negative or huge values and for NaN the division
// decides.  Same result as axis_index for every input.
__host__ __device__ __forceinline__ bool axis_index_fast(double p, double mn, double dx, double inv_dx, uint32_t n,
                                                uint32_t& out) {
    const double y = (p - mn) * inv_dx;
    const double fl = floor(y);
    const double fr = y - fl;
    if (y >= 0.0 && y < 4294967296.0 && fr > 1e-4 && fr < 1.0 - 1e-4) {
        const uint32_t i = (uint32_t)fl;
        if (i >= n) return false;
        out = i;
        return true;
    }
    return axis_index(p, mn, dx, n, out);
}

// a / d for a, d < 2^32 with inv_d = 1 / d in double: the estimate is off by at
// most one, which one remainder test corrects.
__host__ __device__ __forceinline__ uint32_t udiv_by_inv(uint32_t a, uint32_t d, double inv_d) {
    uint32_t q = (uint32_t)((double)a * inv_d);
    const int64_t r = (int64_t)a - (int64_t)q * (int64_t)d;
    if (r < 0) q -= 1u;
    else if (r >= (int64_t)d) q += 1u;
    return q;
}

}  // namespace crm
