// lima_terms.h -- LIMA's bonded-term and PME-stencil arithmetic as device functions, shared by the
// per-step engine kernels and the batched minimiser so both evaluate identical formulas.
#pragma once

#include "lima_dev.h"

namespace lima {

// Bonded-term incidence codes: (type << 29) | (role << 27) | term.  An atom's incidence list holds
// one code per term it belongs to; role is its position in the term (0 .. 3).
enum { T_BOND = 0, T_ANGLE = 1, T_DIH = 2, T_IMP = 3, T_PAIR = 4 };
constexpr int INC_MAX_TERMS = 1 << 27;
__host__ __device__ inline uint32_t inc_code(int type, int role, int term) {
    return ((uint32_t)type << 29) | ((uint32_t)role << 27) | (uint32_t)term;
}

struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ float len(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ v3 nrm(v3 a) { return a * (1.0f / len(a)); }

// A displacement in integer Coord units (1e-8 nm), exact.
struct Dq {
    long long x, y, z;
};
__device__ __forceinline__ v3 dq_nm(const Dq& d) {
    return {(float)d.x * LIMA_LIMA_TO_NANO_F, (float)d.y * LIMA_LIMA_TO_NANO_F, (float)d.z * LIMA_LIMA_TO_NANO_F};
}

// |d| - r0 in Coord units, and |d| (*len_q), for a bond-like length close to its rest length r0 [nm].
// Rounding |d| to FP32 costs about 6e-8 |d|, i.e. ~1e-5 of the deviation when a bond sits 1e-3 nm
// from r0.  Here |d|^2 is summed exactly in integers and split into two floats, the square root
// is corrected by its exact FMA residual, and r0 * 1e8 is an exact two-float product (1e8 is a
// float), so the deviation keeps FP32 precision of its own size.
__device__ __forceinline__ float length_deviation_q(const Dq& d, float r0_nm, float* len_q) {
#pragma clang fp contract(off)
    const long long r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const float hi = (float)r2, lo = (float)(r2 - (long long)hi);
    const float rh = sqrtf(hi);
    const float rl = (fmaf(-rh, rh, hi) + lo) / (2.f * rh);
    const float bh = r0_nm * 1e8f, bl = fmaf(r0_nm, 1e8f, -bh);
    *len_q = rh;
    return (rh - bh) + (rl - bl);
}

// ---- bonded terms (ForceComputations.cuh).  Each returns the term energy [J/mol] and the
// forces [J/mol/nm] on its atoms; arguments are displacements in nm. ----

// calcSinglebondForces, ForceComputations.cuh:17-51.  d = x_a - x_b in Coord units; p = (b0, kb).
// Force on a is *fa, on b -*fa.
__device__ __forceinline__ float term_bond(const Dq& d, float2 p, v3* fa) {
    float rq;
    const float err = length_deviation_q(d, p.x, &rq) * LIMA_LIMA_TO_NANO_F;
    const float s = -p.y * err / rq;   // force along the unit vector d / |d|
    *fa = {(float)d.x * s, (float)d.y * s, (float)d.z * s};
    return 0.5f * p.y * err * err;
}

// calcAnglebondForces, ForceComputations.cuh:53-115: harmonic in theta plus Urey-Bradley on r13.
// r1 = x_l - x_m, r2 = x_r - x_m [nm], d13 = x_l - x_r [Coord units]; p = (theta0, ktheta, ub0, kub).
// Forces on l and r; the middle atom takes -(fl + fr).
__device__ __forceinline__ float term_angle(v3 r1, v3 r2, const Dq& d13, float4 p, v3* fl, v3* fr) {
    const float l1 = len(r1), l2 = len(r2);
    const v3 v1 = r1 * (1.f / l1), v2 = r2 * (1.f / l2);
    const float c = fminf(1.f, fmaxf(-1.f, dot(v1, v2)));
    const float th = acosf(c), err = th - p.x;
    const float s = fmaxf(sqrtf(fmaxf(0.f, 1.f - c * c)), 1e-6f);
    const float tq = p.y * err;
    v3 a = (v2 - v1 * c) * (tq / (l1 * s));
    v3 b = (v1 - v2 * c) * (tq / (l2 * s));
    float e = 0.5f * p.y * err * err;
    if (p.w != 0.f) {
        float r13q;
        const float eu = length_deviation_q(d13, p.z, &r13q) * LIMA_LIMA_TO_NANO_F;
        const float su = -p.w * eu / r13q;
        const v3 fu = {(float)d13.x * su, (float)d13.y * su, (float)d13.z * su};
        a = a + fu;
        b = b - fu;
        e += 0.5f * p.w * eu * eu;
    }
    *fl = a;
    *fr = b;
    return e;
}

// calcDihedralbondForces, ForceComputations.cuh:119-210: V = k (1 + cos(n phi_L - phi0)) with
// phi_L = -atan2(sin, cos), as the exact gradient of that energy.  b1 = x1 - x0, b2 = x2 - x1,
// b3 = x3 - x2; p = (phi0, k, n).
__device__ __forceinline__ float term_dihedral(v3 b1, v3 b2, v3 b3, float4 p, v3* f) {
    const v3 m = cross(b1, b2), n = cross(b2, b3);
    const float lb2 = len(b2);
    const float phi = -atan2f(lb2 * dot(b1, n), dot(m, n));
    const float arg = p.z * phi - p.x;
    const float dv = -p.y * p.z * sinf(arg);   // dV/dphi_L
    const v3 g1 = m * (-lb2 / dot(m, m));
    const v3 g4 = n * (lb2 / dot(n, n));
    const float pp = dot(b1, b2) / (lb2 * lb2), qq = dot(b3, b2) / (lb2 * lb2);
    const v3 g2 = g1 * (-(pp + 1.f)) + g4 * qq;
    const v3 g3 = g1 * pp - g4 * (qq + 1.f);
    f[0] = g1 * dv;
    f[1] = g2 * dv;
    f[2] = g3 * dv;
    f[3] = g4 * dv;
    return p.y * (1.f + cosf(arg));
}

// calcImproperdihedralbondForces, ForceComputations.cuh:214-264 (LIMA's force split, not a
// gradient).  pj, pk, pl are x_j - x_i, x_k - x_i, x_l - x_i; p = (psi0, kpsi).
__device__ __forceinline__ float term_improper(v3 pj, v3 pk, v3 pl, float2 p, v3* f) {
    const v3 pi = {0.f, 0.f, 0.f};
    const v3 ij = nrm(pj - pi), ik = nrm(pk - pi), il = nrm(pl - pi), lj = nrm(pj - pl), lk = nrm(pk - pl);
    const v3 n1 = nrm(cross(ij, ik)), n2 = nrm(cross(lj, lk));
    // the angle between the plane normals (LIMA: acos of their dot product), as atan2 of sine and
    // cosine: acos loses ~1e-3 of a 0.01 rad angle in FP32, and impropers sit near their minimum
    float ang = atan2f(len(cross(n1, n2)), dot(n1, n2));
    if (dot(n1, il) > 0.f) ang = -ang;
    const float err = ang - p.x;
    const float tq = p.y * err;
    const float dl = len(cross(pk - pj, pj - pl)) / len(pk - pj);
    const float di = len(cross(pk - pj, pj - pi)) / len(pk - pj);
    const v3 fl = n2 * (tq / dl);
    const v3 fi = -n1 * (tq / di);
    const v3 res = -(fi + fl);
    const float ratio = len(pj - pi) / (len(pj - pi) + len(pk - pi));
    f[0] = fi;
    f[1] = res * (1.f - ratio);
    f[2] = res * ratio;
    f[3] = fl;
    return 0.5f * p.y * err * err;
}

// computePairbondForces, ForceComputations.cuh:335-376: 1-4 LJ with the pair's own sigma / eps,
// no Coulomb.  d = x_b - x_a; p = (sigma, eps).  Force on a is *fa, on b -*fa.
__device__ __forceinline__ float term_pair14(v3 d, float2 p, v3* fa) {
    const float r2 = dot(d, d), rinv2 = 1.f / r2;
    float s = p.x * p.x * rinv2;
    s = s * s * s;
    *fa = d * (24.f * p.y * s * rinv2 * (1.f - 2.f * s));
    return 4.f * p.y * s * (s - 1.f);
}

// ---- PME stencil (PME.cuh:263-300, DeviceAlgorithmsPrivate.cuh:7-13) ----

// Grid coordinate of a wrapped integer Coord w in [0, Lq): base node ix = floor(w K / Lq) and
// fraction f, exact in integers.
__device__ __forceinline__ void grid_coord(uint32_t w, long long Lq, int K, int* ix, float* f) {
    // ix = floor(w K / Lq) exactly: w K < 2^41 and Lq are exact doubles, so the double product with
    // 1/Lq is within one unit of the floor, and one integer correction step makes it exact
    const long long num = (long long)((unsigned long long)w * (unsigned long long)K);
    long long q = (long long)((double)num * (1.0 / (double)Lq));
    long long r = num - q * Lq;
    if (r < 0) { q--; r += Lq; }
    else if (r >= Lq) { q++; r -= Lq; }
    *ix = (int)q;
    *f = (float)r / (float)Lq;
}

// f[role] for the 4 per-atom forces of a dihedral or improper, as a sum weighted by 0 / 1 so the
// array stays in registers (a select between elements is folded into an indexed load, which puts
// the array in scratch memory).
__device__ __forceinline__ v3 pick_role4(const v3* f, int role) {
    const float w0 = role == 0 ? 1.f : 0.f, w1 = role == 1 ? 1.f : 0.f, w2 = role == 2 ? 1.f : 0.f, w3 = role == 3 ? 1.f : 0.f;
    v3 o;
    o.x = fmaf(f[3].x, w3, fmaf(f[2].x, w2, fmaf(f[1].x, w1, f[0].x * w0)));
    o.y = fmaf(f[3].y, w3, fmaf(f[2].y, w2, fmaf(f[1].y, w1, f[0].y * w0)));
    o.z = fmaf(f[3].z, w3, fmaf(f[2].z, w2, fmaf(f[1].z, w1, f[0].z * w0)));
    return o;
}

// PME tile bins of charged particles (tiled spreading).  A particle goes to the fixed-capacity
// bin of the tile holding its base node; the bin order does not matter (the tile sums are
// integer).  The counters start at zero each evaluation: the tiled gather clears them after use.
// bin_count == nullptr: no tiled PME, nothing to bin.
struct BinArgs {
    long long Lq[3];
    int K[3], T[3], nt[3];
    int bincap;
    int* bin_count;
    int* bin_slot;     // [tile][bincap] the particle's slot (where the gather writes its force)
    uint4* bin_pos;    // [tile][bincap] its wrapped Coord and charge bits: the spread and gather read
                       // their tile's particles contiguously instead of gathering them by index
    int* flags;        // bit 0: a bin overflowed
};

// Called by all 64 lanes of a wave (converged); lanes with active set bin the particle of slot s at
// pw with charge q.
// Neighbouring lanes mostly share a tile (slots follow the cluster order): ballots group the
// lanes by tile (leader = lowest lane of the group, rank = lanes of the group below), then every
// leader reserves its group's range with one atomic, all in the same instruction (one round trip
// per wave), and each lane takes its base from its leader.
__device__ __forceinline__ void bin_wave(const BinArgs& b, bool active, int s, uint4 pw, float q) {
    int tile = -1;
    if (active) {
        int ix[3];
        float f;
        grid_coord(pw.x, b.Lq[0], b.K[0], &ix[0], &f);
        grid_coord(pw.y, b.Lq[1], b.K[1], &ix[1], &f);
        grid_coord(pw.z, b.Lq[2], b.K[2], &ix[2], &f);
        tile = ((ix[2] / b.T[2]) * b.nt[1] + ix[1] / b.T[1]) * b.nt[0] + ix[0] / b.T[0];
    }
    const int lane = threadIdx.x & 63;
    int leader = lane, rank = 0, cnt = 0;
    unsigned long long pending = __ballot(active);
    while (pending) {   // wave-uniform: one pass per distinct tile, registers only
        const int L = __builtin_ctzll(pending);
        const int t = __shfl(tile, L, 64);
        const bool mine = active && tile == t;
        const unsigned long long same = __ballot(mine);
        if (mine) {
            leader = L;
            rank = __popcll(same & ((1ull << lane) - 1ull));
        }
        if (lane == L) cnt = __popcll(same);
        pending &= ~same;
    }
    int base = 0;
    if (active && lane == leader) base = atomicAdd(&b.bin_count[tile], cnt);
    base = __shfl(base, leader, 64);
    if (active) {
        const int pos = base + rank;
        if (pos < b.bincap) {
            const int64_t e = (int64_t)tile * b.bincap + pos;
            b.bin_slot[e] = s;
            b.bin_pos[e] = make_uint4(pw.x, pw.y, pw.z, __float_as_uint(q));
        } else {
            atomicOr(b.flags, 1);
        }
    }
}

// Cubic B-spline weights at nodes ix-1 .. ix+2 (DeviceAlgorithmsPrivate.cuh:9-12).
__device__ __forceinline__ void bspline(float f, float* w) {
    const float g = 1.f - f;
    w[0] = g * g * g * (1.f / 6.f);
    w[1] = (4.f - 6.f * f * f + 3.f * f * f * f) * (1.f / 6.f);
    w[2] = (1.f + 3.f * f + 3.f * f * f - 3.f * f * f * f) * (1.f / 6.f);
    w[3] = f * f * f * (1.f / 6.f);
}

__device__ __forceinline__ int wrapi(int v, int K) { return v < 0 ? v + K : (v >= K ? v - K : v); }

// Gather weights over the 6 nodes ix-2 .. ix+3 of one axis: w6 interpolates phi, d6 is the
// central-difference stencil of the field (PME.cuh:374-411) folded onto the same nodes:
// -sum_n w(n) (phi(n+1) - phi(n-1)) = sum_m phi(m) (w(m+1) - w(m-1)).
__device__ __forceinline__ void gather_weights(const float* w, float* w6, float* d6) {
    for (int m = 0; m < 6; m++) {
        const int k = m - 1;
        w6[m] = (k >= 0 && k < 4) ? w[k] : 0.f;
        const int lo = m - 2, hi = m;
        d6[m] = ((hi >= 0 && hi < 4) ? w[hi] : 0.f) - ((lo >= 0 && lo < 4) ? w[lo] : 0.f);
    }
}

}  // namespace lima
