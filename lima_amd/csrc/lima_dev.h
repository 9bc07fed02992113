// lima_dev.h -- device-side helpers shared by the CDNA4 kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// ---- LIMA constants (code/LIMA_BASE/include/Constants.h:43-66, PhysicsUtils.cuh:5) ----
#define LIMA_PI_F 3.14159f
#define LIMA_NANO_TO_LIMA_F 100000000.0f
#define LIMA_LIMA_TO_NANO_F 1e-8f

namespace lima {

constexpr int WAVE = 64;
constexpr int CLUSTER = 8;          // particles per cluster (pair-list granularity)
constexpr float PAD_POS = 1.0e4f;   // position of padding slots, far outside every cutoff

struct f3 { float x, y, z; };

// Image shift of one axis, code 0..2 for -1, 0, +1 box lengths; codes of a list entry are
// sx + 3 sy + 9 sz.  shift_code_axes splits a code < 27 without a division.
__host__ __device__ __forceinline__ long long shift_of(int c, long long Lq) { return c == 0 ? -Lq : (c == 2 ? Lq : 0ll); }
__device__ __forceinline__ void shift_code_axes(int cd, int* cx, int* cy, int* cz) {
    const int z = (cd * 57) >> 9, rem = cd - 9 * z, y = (rem * 11) >> 5;
    *cz = z; *cy = y; *cx = rem - 3 * y;
}


__device__ __forceinline__ int64_t unwrapped(uint32_t w, int32_t img, int64_t Lq) { return (int64_t)w + (int64_t)img * Lq; }

// Per-slot force records (fx, fy, fz, potential-energy share).  The energy share is written only by
// energy evaluations: the other evaluations store 12 bytes and leave it as the last energy
// evaluation wrote it.
template <bool ENERGY>
__device__ __forceinline__ void put_force(float4* buf, int s, float fx, float fy, float fz, float e) {
    if (ENERGY) {
        buf[s] = make_float4(fx, fy, fz, e);
    } else {
        float* p = reinterpret_cast<float*>(buf + s);
        p[0] = fx; p[1] = fy; p[2] = fz;
    }
}

// Block index remap: workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Block b runs on XCD b mod 8; the remap gives each XCD a contiguous range of logical blocks
// (a bijection on [0, nb)), so neighbouring logical blocks, which share data, share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
// The same for a 2D grid: (x, y) of the logical block, x fastest.
__device__ __forceinline__ int2 xcd_block2() {
    const int nb = gridDim.x * gridDim.y;
    const int l = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, nb);
    const int y = l / gridDim.x;
    return make_int2(l - y * gridDim.x, y);
}

// Sum over aligned groups of 8 lanes: quad_perm xor1, quad_perm xor2, row_half_mirror.
// Every lane of the group ends up with the group sum.
__device__ __forceinline__ float sum8(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    return v;
}

// Wave-wide sum (any lane count), result valid in every lane.
__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// float atomic max for non-negative values (IEEE order == integer order)
__device__ __forceinline__ void atomic_max_pos(float* addr, float v) {
    atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

// LIMA pair interaction in FP32 (LennardJonesInteractions.cuh:74-108,
// DeviceAlgorithmsPrivate.cuh:38-95): returns S with force on self = diff * S,
// diff = x_other - x_self.  sig = sigma_i/2 + sigma_j/2, eps = sqrt(eps_i) sqrt(eps_j).
struct PairConst {
    float kc;          // modifiedCoulombConstant
    float kappa;       // 3 / cutoff
    float two_kappa_sqrtpi;   // 2 kappa / sqrt(PI_L)
};

__device__ __forceinline__ float pair_scalar(float r2, float sig, float eps, float qq, const PairConst& c,
                                             float* e_lj, float* e_c) {
    const float rinv2 = 1.0f / r2;
    float s = sig * sig * rinv2;
    s = s * s * s;
    const float flj = 24.0f * eps * s * rinv2 * (1.0f - 2.0f * s);
    const float r = sqrtf(r2);
    const float kr = c.kappa * r;
    const float ec = erfcf(kr);
    const float g = ec + c.two_kappa_sqrtpi * r * expf(-kr * kr);
    const float fc = c.kc * qq * rinv2 / r * g;
    if (e_lj) {
        *e_lj = 4.0f * eps * s * (s - 1.0f);
        *e_c = c.kc * qq * ec / r;
    }
    return flj - fc;
}

// The same pair interaction for kappa r <= 3.25 (every pair inside the cutoff, kappa = 3 / cutoff),
// with hardware rsq / rcp / exp2 instead of IEEE divisions and libm calls:
//   erfc(x) = exp(-x^2) R(x),  R(x) = erfc(x) exp(x^2) ~= sum_{k=1..7} a_k t^k,  t = 1 / (1 + 0.35 x).
// The a_k are a relative-error least-squares fit of R on [0, 3.25] (max relative error 1e-7 in
// exact arithmetic, 4e-7 absolute on erfc when evaluated in FP32).  exp(-x^2) serves both the
// erfc and the Gaussian term of the force.  eps24 = 24 eps.
__device__ __forceinline__ float pair_scalar_cut(float r2, float sig, float eps24, float qq_kc, const PairConst& c,
                                                 float* e_lj, float* e_c) {
    // qq_kc = kC q_i q_j; exp(-x^2) = exp2(r^2 (-kappa^2 log2 e)), the constant hoisted by the caller's loop
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    float s = sig * sig * rinv2;
    s = s * s * s;
    const float flj = eps24 * s * rinv2 * (1.0f - 2.0f * s);
    const float ex = __builtin_amdgcn_exp2f(r2 * (-(c.kappa * c.kappa) * 1.4426950408889634f));
    const float t = __builtin_amdgcn_rcpf(1.0f + (0.35f * c.kappa) * (r2 * rinv));   // 1 / (1 + 0.35 kappa r)
    float R = -2.035868019e-01f;
    R = R * t + 7.992476821e-01f;
    R = R * t - 7.338286042e-01f;
    R = R * t + 7.401084900e-01f;
    R = R * t - 4.074399173e-02f;
    R = R * t + 2.456984967e-01f;
    R = R * t + 1.931048036e-01f;
    const float Rt = R * t;   // erfc(x) = ex Rt
    const float fc = (qq_kc * rinv2) * (ex * fmaf(rinv, Rt, c.two_kappa_sqrtpi));
    if (e_lj) {
        *e_lj = (eps24 * (1.0f / 6.0f)) * s * (s - 1.0f);
        *e_c = qq_kc * (ex * Rt) * rinv;
    }
    return flj - fc;
}

// Coulomb force factor H(r) = erfc(kappa r) + (2 kappa / sqrt(PI_L)) r exp(-kappa^2 r^2) of the
// pair force above (force on self = diff * kC qq H(r) / r^3), tabulated as a function of z = r^2
// on [0, cutoff^2] in COUL_TAB intervals as cubic Hermite polynomials in the local coordinate
// t in [0, 1) (c0 + c1 t + c2 t^2 + c3 t^3; values and z-derivatives of H in double at the nodes),
// plus one entry for z = cutoff^2 itself.  H(sqrt z) = 1 - (4 kappa^3 / 3 sqrt(pi)) z^1.5 + ...
// is smooth enough in z that the interpolation error stays below 3e-7 of H for r >= 0.1 nm
// (h = cutoff^2 / 256); indexing by r^2 saves the square root and the index clamp.
constexpr int COUL_TAB = 256;

// pair_scalar_cut with the Coulomb force factor read from the LDS table (tab[COUL_TAB + 1], zscale =
// COUL_TAB / cutoff^2); r2 <= cutoff^2.  The energies, when asked for, are evaluated as in
// pair_scalar_cut.
__device__ __forceinline__ float pair_scalar_tab(float r2, float sig, float eps24, float qq_kc, const PairConst& c,
                                                 const float4* tab, float zscale, float* e_lj, float* e_c) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    float s = sig * sig * rinv2;
    s = s * s * s;
    const float x = r2 * zscale;
    const float4 k = tab[(int)x];
    const float t = __builtin_amdgcn_fractf(x);
    const float H = fmaf(fmaf(fmaf(k.w, t, k.z), t, k.y), t, k.x);
    // LJ eps24 s rinv2 (1 - 2 s) minus Coulomb qq H rinv^3, with the common rinv2 factored out
    const float S = rinv2 * fmaf(eps24 * s, fmaf(-2.f, s, 1.f), -((qq_kc * rinv) * H));
    if (e_lj) {
        *e_lj = (eps24 * (1.0f / 6.0f)) * s * (s - 1.0f);
        float el, ecl;
        pair_scalar_cut(r2, sig, eps24, qq_kc, c, &el, &ecl);
        *e_c = ecl;
    }
    return S;
}

}  // namespace lima
