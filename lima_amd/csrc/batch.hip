// batch.hip -- batched steepest-descent minimisation of many small periodic systems in one
// persistent launch per workgroup shape.
//
// Each workgroup owns one system and runs its whole minimisation (Programs::EnergyMinimize,
// Programs.cpp:155-199, with steepest descent and FP64 acceptance instead of LIMA's ADAM EM):
//
//   E, F <- forces(x); h <- h0
//   while Fmax >= tol and iterations < max:
//       x' = x + trunc(h F / Fmax)            (integer Coord, per atom)
//       E', F' <- forces(x')
//       accept (x, F, E <- x', F', E'; h *= 1.2) if E' < E in FP64, else h *= 0.2
//
// Force evaluation inside the workgroup, with no inter-workgroup synchronisation:
//   - nonbonded cutoff pairs (LennardJonesInteractions.cuh:74-203): every pair of the system once,
//     as pairs of 64-atom tiles with the j tile rotating across the wave; boxes are 1.5-3.5 nm
//     against a 1.2 nm cutoff, so the cutoff sphere covers 25-100 % of the box and a cell list
//     would not prune.
//     Every periodic image within the cutoff counts (two candidates per axis when an edge is
//     shorter than 2 rc).  Pairs inside one compound or between bonded compounds are skipped
//     here and computed by the intra pass.
//   - intra pairs (LennardJonesInteractions.cuh:121-161): same / bonded compound, BondedParticlesLUT
//     exclusions, no cutoff, unwrapped displacement.
//   - bonded terms (ForceComputations.cuh): each atom sums the terms it belongs to from a per-atom
//     incidence list (the term is evaluated once per member atom, so no atomics).
//   - PME (PME.cuh): int32 fixed-point spreading into an LDS grid, DFT passes on the per-system
//     half spectrum, LIMA's influence function, finite-difference gather from an LDS potential.
//   Energies are reduced in FP64 in a fixed order; the run is deterministic.
//
// Block shapes (threads per workgroup, unroll of the pair loop) come from lima_batch_autotune,
// which times every candidate shape on the batch's own systems.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "lima_terms.h"
#include "lima_mlp.h"

namespace lima {
namespace {

constexpr int KMAX = 32;                       // largest PME grid edge (twiddle tables, LDS grid)
constexpr int MAX_ATOMS = 8192;                // atoms per system
constexpr float FIX_SCALE = 1048576.0f;        // 2^20 fixed-point units per kC/mol of spread charge
constexpr int NBUCKET = 4;
constexpr int BUCKET_MAX[NBUCKET] = {600, 1200, 2000, MAX_ATOMS};

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

struct SysMeta {
    int n, atom0, nc, comp0;
    int Lq[3], hLq[3];
    float L[3];
    int K[3], H, NG;
    int spec0;          // offset of this system's half spectrum and influence function
    int part0;          // offset (float4) of the tile-pair partial forces: [NT(NT+1)/2][2][64]
    int tw0[3];         // twiddle table offsets per axis
    float fix_to_rho;   // fixed point -> charge density: 1 / (2^20 vcell)
    float inv_ng;
    float inv2h[3];
    double eself;       // -kappa / sqrt(pi_L) k_C sum q^2 (standard Ewald self energy)
    int multi;          // an edge shorter than 2 rc: up to two images per axis
    int elec;           // system has charges and electrostatics on
    int lds_bytes;
};

struct SysOut {
    double e_acc;       // accepted potential energy [J/mol], incl. the Ewald self energy
    double e_parts[4];  // lj, coulomb short range, bonded, pme reciprocal of the accepted state
    float fmax;         // [J/mol/nm]
    float h;
    int iterations, accepted, converged, cur;
    int evaluations;             // force evaluations (learned mode: up to two per iteration)
    int learned_accepted;        // iterations whose learned proposal was accepted
    int rec_iters;               // recorded states (recording mode)
    unsigned long long t0, t1;   // s_memrealtime (100 MHz) at start / end
    unsigned long long ph[7];    // ticks in tile pairs, per-atom sums + intra + bonded, PME, reductions, trial steps,
                                 // feature histograms, scale inference
};

struct BatchArgs {
    const SysMeta* meta;
    const int* order;            // system of each workgroup
    int4* u0; int4* u1;          // unwrapped Coord (two buffers: accepted / trial)
    float4* f0; float4* f1;      // forces
    const float4* prm;           // (q kC/mol, sigma/2 nm, sqrt(eps), sqrt(24 eps))
    const int* comp;             // local compound of each atom
    const int4* cinfo;           // per compound (first atom local, atom count, key atom local)
    const int* cbm;              // per compound bonded-compound mask (bit k: compound c+k-2)
    const uint32_t* clut;        // per compound [5][32] exclusion rows
    const int* inc_off;          // per atom incidence offsets (global atom index, N+1)
    const int* inc;              // (type << 29) | (role << 27) | term
    const int2* bid; const float2* bp;
    const int4* aid; const float4* ap;
    const int4* did; const float4* dp;
    const int4* iid; const float2* ip;
    const int2* pid; const float2* pp;
    const float* green;
    float2* spec;
    float2* spec2;               // second half-spectrum buffer (the transforms ping-pong)
    float4* part;
    const float2* twid;
    const float4* ctab;          // Coulomb factors of the cutoff pairs in r (see CTAB_N), staged in LDS
    float ctab_scale;            // CTAB_N / cutoff
    SysOut* out;
    PairConst pc;
    float rc, rc2;
    double tol;                  // Fmax tolerance [J/mol/nm]
    double h0;                   // initial step [nm]
    int max_iter;
    int mode;                    // MODE_PLAIN, MODE_LEARNED, MODE_RECORD
    MlpWeights mlp;              // learned mode: step-scale model
    unsigned short* feat;        // [N][32] BF16 features (slots 4..31: neighbour histogram)
    float* scale;                // [N] learned step scale of the current state
    int feat_interval;           // iterations between histogram refreshes (learned mode)
    float* rec;                  // recording: [rec_info.x + (state * nsamp + k) * REC_W]
    const int2* rec_info;        // per system (offset in floats, sampled atoms)
    int rec_stride, rec_cap;     // sampled atoms li = k * rec_stride; states per system
    int grow_on_learned;         // learned mode: h *= 1.2 after an accepted proposal (else h is kept)
    int wlds;                    // learned mode: byte offset of the BF16 weights in dynamic LDS, staged once
    unsigned* hnext;             // learned mode: [N][7] packed neighbour-histogram bytes of the last trial
                                 // evaluated before a feature refresh (counted in its pair pass)
};

enum { MODE_PLAIN = 0, MODE_LEARNED = 1, MODE_RECORD = 2 };
constexpr int REC_W = 48;              // floats per record: 32 features, step (3), Coord (3), force (3), pad
constexpr float HIST_R = 1.0f;         // neighbour histogram radius [nm]
constexpr int HIST_BINS = 14;          // bins per charge sign
constexpr int HIST_MAX_ATOMS = 4096;   // byte counters in LDS: 28 per atom
// Cutoff pairs read erfc(kappa r) and H(r) = erfc(kappa r) + (2 kappa / sqrt(PI_L)) r exp(-kappa^2 r^2)
// from a table in r: CTAB_N intervals on [0, cutoff], per interval two cubic Hermite polynomials in
// the local coordinate t in [0, 1) (erfc, then H; values and derivatives in double at the nodes),
// plus the values at r = cutoff.  Both are analytic in r, so the interpolation error is ~1e-10.
constexpr int CTAB_N = 256;


__device__ __forceinline__ int fmodi(int x, int L) {
    int r = x % L;
    return r < 0 ? r + L : r;
}

__device__ __forceinline__ int min_image(int d, int Lq) {   // representative of d mod Lq in [-Lq/2, Lq/2]
    int r = d % Lq;
    r += r < 0 ? Lq : 0;
    r -= 2 * r > Lq ? Lq : 0;
    return r;
}

// x_b - x_a [nm] of bonded particles: the nearest image (LIMA's hyperpos result for particles a
// few bonds apart; the Coords of different compounds may differ by whole boxes)
__device__ __forceinline__ Dq bdispq(const int4* u, int a, int b, const SysMeta& m) {
    const int4 ua = u[a], ub = u[b];
    return {(long long)min_image(ub.x - ua.x, m.Lq[0]), (long long)min_image(ub.y - ua.y, m.Lq[1]),
            (long long)min_image(ub.z - ua.z, m.Lq[2])};
}
__device__ __forceinline__ v3 bdisp(const int4* u, int a, int b, const SysMeta& m) { return dq_nm(bdispq(u, a, b, m)); }

// pair_scalar_cut with the Coulomb factors read from the LDS table in r (r2 <= cutoff^2):
// force scalar S (force on i = d S) and the LJ and Coulomb energies of the pair.
__device__ __forceinline__ float pair_scalar_rtab(float r2, float sig, float eps24, float qq_kc, const float4* __restrict__ tab,
                                                  float rscale, float* e_lj, float* e_c) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    float s = sig * sig * rinv2;
    s = s * s * s;
    const float x = (r2 * rinv) * rscale;
    const int k = (int)x;
    const float t = __builtin_amdgcn_fractf(x);
    const float4 ka = tab[2 * k], kb = tab[2 * k + 1];
    const float A = fmaf(fmaf(fmaf(ka.w, t, ka.z), t, ka.y), t, ka.x);   // erfc(kappa r)
    const float H = fmaf(fmaf(fmaf(kb.w, t, kb.z), t, kb.y), t, kb.x);
    const float qr = qq_kc * rinv;
    *e_lj = (eps24 * (1.0f / 6.0f)) * s * (s - 1.0f);
    *e_c = qr * A;
    return rinv2 * fmaf(eps24 * s, fmaf(-2.f, s, 1.f), -(qr * H));
}

// One cutoff pair computed once: force on i += d S, on j -= d S, the whole pair energy.
// pj = (q, sigma/2, sqrt(24 eps)) of j.
__device__ __forceinline__ void pair_eval(float x, float y, float z, float r2, const float4& pi, float qj, float sj, float ej,
                                          const PairConst& pc, const float4* tab, float rscale, v3& fi, v3& fj, float& elj, float& ec) {
    float e1, e2;
    const float S = pair_scalar_rtab(r2, pi.y + sj, pi.w * ej, pi.x * qj * pc.kc, tab, rscale, &e1, &e2);
    // explicit FMAs: the same rounding whatever the unrolling or scheduling
    fi.x = fmaf(x, S, fi.x); fi.y = fmaf(y, S, fi.y); fi.z = fmaf(z, S, fi.z);
    fj.x = fmaf(-x, S, fj.x); fj.y = fmaf(-y, S, fj.y); fj.z = fmaf(-z, S, fj.z);
    elj += e1;
    ec += e2;
}

// Box constants of one system held in registers for the pair loop.
struct PairBox {
    int Lx, Ly, Lz, hx, hy, hz;
    float fx, fy, fz;      // box edges [nm]
    float ix, iy, iz;      // their reciprocals
};

// Cutoff pair (i, j) over every periodic image within the cutoff.  `excl` marks a pair of one
// compound or of bonded compounds (left to the intra pass).  One predicate guards the pair
// arithmetic, so a wave branches once per step.
// Neighbour-histogram counting folded into a pair pass (learned mode): byte counters in LDS, bin b
// of the distance, offset by HIST_BINS when the partner is negative; every pair within HIST_R is
// counted, excluded (intra / bonded) pairs included, once per image.
struct HistCtx {
    unsigned* hc;     // [n][7] packed byte counters
    int li, lj;       // local atoms of the pair
    int ineg, jneg;   // HIST_BINS when the atom's charge is negative, else 0
    bool on;          // counting in this pass (wave-uniform)
    bool skip;        // self pair, padding or the pair's second visit in a diagonal tile
};
__device__ __forceinline__ void hist_count(const HistCtx& h, float r2) {
    if (r2 <= HIST_R * HIST_R && !h.skip) {
        const int b = min(HIST_BINS - 1, (int)(sqrtf(r2) * ((float)HIST_BINS / HIST_R)));
        const int bi = b + h.jneg, bj = b + h.ineg;
        atomicAdd(&h.hc[h.li * 7 + (bi >> 2)], 1u << ((bi & 3) * 8));
        atomicAdd(&h.hc[h.lj * 7 + (bj >> 2)], 1u << ((bj & 3) * 8));
    }
}

template <bool MULTI>
// Positions are the wrapped Coord in nm as floats (rounding ~2e-7 nm at these box sizes); the
// nearest image is d - L rint(d / L), four instructions per axis.
__device__ __forceinline__ void pair_cut(const float3& wi, const float4& pi, bool excl, float jx, float jy, float jz, float qj,
                                         float sj, float ej, const PairBox& bx, const BatchArgs& a, const float4* tab, v3& fi, v3& fj,
                                         float& elj, float& ec, const HistCtx& hist) {
    float x = jx - wi.x, y = jy - wi.y, z = jz - wi.z;
    x = fmaf(-bx.fx, rintf(x * bx.ix), x);
    y = fmaf(-bx.fy, rintf(y * bx.iy), y);
    z = fmaf(-bx.fz, rintf(z * bx.iz), z);
    if (!MULTI) {
        const float r2 = x * x + y * y + z * z;
        if (hist.on) hist_count(hist, r2);
        if (!excl && r2 <= a.rc2) pair_eval(x, y, z, r2, pi, qj, sj, ej, a.pc, tab, a.ctab_scale, fi, fj, elj, ec);
    } else if (!excl || hist.on) {
        // second candidate image per axis: the one on the other side, when within the cutoff
        const float ax = x - copysignf(bx.fx, x), ay = y - copysignf(bx.fy, y), az = z - copysignf(bx.fz, z);
        const int nx = fabsf(ax) <= a.rc ? 2 : 1, ny = fabsf(ay) <= a.rc ? 2 : 1, nz = fabsf(az) <= a.rc ? 2 : 1;
        for (int kz = 0; kz < nz; kz++) {
            const float zz = kz ? az : z;
            for (int ky = 0; ky < ny; ky++) {
                const float yy = ky ? ay : y;
                for (int kx = 0; kx < nx; kx++) {
                    const float xx = kx ? ax : x;
                    const float r2 = xx * xx + yy * yy + zz * zz;
                    if (hist.on) hist_count(hist, r2);
                    if (!excl && r2 <= a.rc2) pair_eval(xx, yy, zz, r2, pi, qj, sj, ej, a.pc, tab, a.ctab_scale, fi, fj, elj, ec);
                }
            }
        }
    }
}

// Rotate a value across the wave by one lane (DPP wave_rol:1): lane l receives lane l+1's value.
__device__ __forceinline__ int rot1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ float rot1f(float v) { return __int_as_float(rot1(__float_as_int(v))); }

// intra pairs of atom i (local li): its compound and the bonded compounds, LUT-excluded, no cutoff
__device__ __forceinline__ void intra_atom(const BatchArgs& a, const SysMeta& m, const int4* u, int li, int ci, int bmi, float4 pi,
                                           v3& acc, float& elj, float& ec) {
    const int gi = m.atom0 + li;
    const int4 ca = a.cinfo[m.comp0 + ci];
    const int si = li - ca.x;
    const int4 ui = u[gi];
    for (int k = 0; k < 5; k++) {
        if (!((bmi >> k) & 1)) continue;
        const int B = ci + k - 2;
        const uint32_t row = a.clut[((size_t)(m.comp0 + ci) * 5 + k) * 32 + si];
        const int4 cb = a.cinfo[m.comp0 + B];
        // within the compound unwrapped; a bonded compound moves to the image whose key particle
        // is nearest to this compound's (LIMA's hyperpos of the neighbour's origo)
        int sx = 0, sy = 0, sz = 0;
        if (B != ci) {
            const int4 ka = u[m.atom0 + ca.z], kb = u[m.atom0 + cb.z];
            sx = min_image(kb.x - ka.x, m.Lq[0]) - (kb.x - ka.x);
            sy = min_image(kb.y - ka.y, m.Lq[1]) - (kb.y - ka.y);
            sz = min_image(kb.z - ka.z, m.Lq[2]) - (kb.z - ka.z);
        }
        for (int sj = 0; sj < cb.y; sj++) {
            if ((row >> sj) & 1u) continue;
            const int gj = m.atom0 + cb.x + sj;
            const int4 uj = u[gj];
            const v3 d = {(float)(uj.x - ui.x + sx) * LIMA_LIMA_TO_NANO_F, (float)(uj.y - ui.y + sy) * LIMA_LIMA_TO_NANO_F,
                          (float)(uj.z - ui.z + sz) * LIMA_LIMA_TO_NANO_F};
            const float r2 = dot(d, d);
            const float4 pj = a.prm[gj];
            float e1, e2;
            // no cutoff: pair_scalar_cut's erfc polynomial stays within 1e-12 absolute beyond kappa r = 3.25
            const float S = pair_scalar_cut(r2, pi.y + pj.y, 24.0f * (pi.z * pj.z), a.pc.kc * (pi.x * pj.x), a.pc, &e1, &e2);
            acc = acc + d * S;
            elj += 0.5f * e1;
            ec += 0.5f * e2;
        }
    }
}

// bonded terms containing atom gi, summed from its incidence list; the term energy is counted by
// the incidence with role 0
__device__ __forceinline__ void bonded_atom(const BatchArgs& a, const SysMeta& m, const int4* u, int gi, v3& acc, float& eb) {
    const int e0 = a.inc_off[gi], e1 = a.inc_off[gi + 1];
    for (int q = e0; q < e1; q++) {
        const int code = a.inc[q];
        const int type = (code >> 29) & 7, role = (code >> 27) & 3, t = code & ((1 << 27) - 1);
        v3 f = {0.f, 0.f, 0.f};
        float e = 0.f;
        if (type == T_BOND) {
            const int2 id = a.bid[t];
            v3 fa;
            e = term_bond(bdispq(u, id.y, id.x, m), a.bp[t], &fa);
            f = role == 0 ? fa : -fa;
        } else if (type == T_ANGLE) {
            const int4 id = a.aid[t];
            v3 fl, fr;
            e = term_angle(bdisp(u, id.y, id.x, m), bdisp(u, id.y, id.z, m), bdispq(u, id.z, id.x, m), a.ap[t], &fl, &fr);
            f = role == 0 ? fl : (role == 2 ? fr : -(fl + fr));
        } else if (type == T_DIH) {
            const int4 id = a.did[t];
            v3 fd[4];
            e = term_dihedral(bdisp(u, id.x, id.y, m), bdisp(u, id.y, id.z, m), bdisp(u, id.z, id.w, m), a.dp[t], fd);
            f = pick_role4(fd, role);
        } else if (type == T_IMP) {
            const int4 id = a.iid[t];
            v3 fd[4];
            e = term_improper(bdisp(u, id.x, id.y, m), bdisp(u, id.x, id.z, m), bdisp(u, id.x, id.w, m), a.ip[t], fd);
            f = pick_role4(fd, role);
        } else {
            const int2 id = a.pid[t];
            v3 fa;
            e = term_pair14(bdisp(u, id.x, id.y, m), a.pp[t], &fa);
            f = role == 0 ? fa : -fa;
        }
        acc = acc + f;
        if (role == 0) eb += e;
    }
}

template <int T>
struct BlockShared {
    double red[T / 64][5];
    float fred[T / 64];
    float fmin_[T / 64];
    double res[5];
    float fres;
    int uniform;                 // infer_scales: every atom got the same scale
    float2 tw[3][KMAX];
    unsigned long long ph[7];    // phase ticks seen by thread 0
    float4 ctab[2 * (CTAB_N + 1)];   // the Coulomb table of the cutoff pairs
};

// FP64 sums of v[0..3] and max of fm over the workgroup, in a fixed order; every thread gets them.
template <int T>
__device__ __forceinline__ void block_reduce(BlockShared<T>& sh, double* v, float fm) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = 0; c < 4; c++) v[c] = wave_sum_d(v[c]);
    fm = wave_max(fm);
    if (lane == 0) {
        for (int c = 0; c < 4; c++) sh.red[wv][c] = v[c];
        sh.fred[wv] = fm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[4] = {0, 0, 0, 0};
        float f = 0.f;
        for (int w = 0; w < T / 64; w++) {
            for (int c = 0; c < 4; c++) s[c] += sh.red[w][c];
            f = fmaxf(f, sh.fred[w]);
        }
        for (int c = 0; c < 4; c++) sh.res[c] = s[c];
        sh.fres = f;
    }
    __syncthreads();
    for (int c = 0; c < 4; c++) v[c] = sh.res[c];
}

// Features 4..31 from packed byte counters: count / 8 as BF16.
template <int T>
__device__ __forceinline__ void hist_features(const BatchArgs& a, const SysMeta& m, const unsigned* hc) {
    for (int k = threadIdx.x; k < m.n; k += T) {
        unsigned short* fo = a.feat + (int64_t)(m.atom0 + k) * 32;
        for (int b = 0; b < 2 * HIST_BINS; b++) {
            const unsigned c = (hc[k * 7 + (b >> 2)] >> ((b & 3) * 8)) & 255u;
            fo[4 + b] = f2bf((float)c * 0.125f);
        }
    }
}

// Neighbour histograms (features 4..31): pairs within HIST_R counted in HIST_BINS distance bins per
// charge sign of the neighbour, every periodic image, each pair once through the same tile-pair walk
// as the forces (j positions rotate across the wave).  Byte counters live in LDS ([n][28]); integer
// increments make the result independent of order.  Written as BF16 count / 8.
template <int T, bool MULTI>
__device__ __forceinline__ void hist_pass(const BatchArgs& a, const SysMeta& m, unsigned char* lds, const int4* u) {
    const int tid = threadIdx.x, n = m.n;
    unsigned* hc = reinterpret_cast<unsigned*>(lds);
    for (int k = tid; k < n * 7; k += T) hc[k] = 0u;
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int W = T / 64;
    const int NT = (n + 63) >> 6, NI = NT * (NT + 1) / 2;
    const PairBox bx = {m.Lq[0], m.Lq[1], m.Lq[2], m.hLq[0], m.hLq[1], m.hLq[2], m.L[0], m.L[1], m.L[2]};
    const float r2max = HIST_R * HIST_R, binv = (float)HIST_BINS / HIST_R;
    const int q0 = (int)((long long)NI * wave / W), q1 = (int)((long long)NI * (wave + 1) / W);
    int I = 0, J = 0;
    {
        int q = q0;
        while (I < NT && q >= NT - I) { q -= NT - I; I++; }
        J = I + q;
    }
    int curI = -1, li = 0, ineg = 0;
    int4 wi = make_int4(0, 0, 0, 0);
    bool vi = false;
    for (int q = q0; q < q1; q++) {
        if (I != curI) {
            curI = I;
            li = I * 64 + lane;
            vi = li < n;
            if (vi) {
                const int4 uu = u[m.atom0 + li];
                wi = make_int4(fmodi(uu.x, bx.Lx), fmodi(uu.y, bx.Ly), fmodi(uu.z, bx.Lz), 0);
                ineg = a.prm[m.atom0 + li].x < 0.f ? HIST_BINS : 0;
            }
        }
        int jx = 0, jy = 0, jz = 0, jl = -1, jneg = 0;
        {
            const int lj = J * 64 + lane;
            if (lj < n) {
                const int4 uu = u[m.atom0 + lj];
                jx = fmodi(uu.x, bx.Lx); jy = fmodi(uu.y, bx.Ly); jz = fmodi(uu.z, bx.Lz);
                jl = lj;
                jneg = a.prm[m.atom0 + lj].x < 0.f ? HIST_BINS : 0;
            }
        }
        const bool diag = I == J;
        const int ns = diag ? 33 : 64;
        for (int st = 0; st < ns; st++) {
            const bool skip = !vi || jl < 0 || (diag && (st == 0 || (st == 32 && lane >= 32)));
            if (!skip) {
                int dx = jx - wi.x, dy = jy - wi.y, dz = jz - wi.z;
                dx += dx < -bx.hx ? bx.Lx : 0; dx -= dx >= bx.hx ? bx.Lx : 0;
                dy += dy < -bx.hy ? bx.Ly : 0; dy -= dy >= bx.hy ? bx.Ly : 0;
                dz += dz < -bx.hz ? bx.Lz : 0; dz -= dz >= bx.hz ? bx.Lz : 0;
                const float x = (float)dx * LIMA_LIMA_TO_NANO_F, y = (float)dy * LIMA_LIMA_TO_NANO_F, z = (float)dz * LIMA_LIMA_TO_NANO_F;
                const float ax = x - copysignf(bx.fx, x), ay = y - copysignf(bx.fy, y), az = z - copysignf(bx.fz, z);
                const int nx = MULTI && fabsf(ax) <= HIST_R ? 2 : 1, ny = MULTI && fabsf(ay) <= HIST_R ? 2 : 1;
                const int nz = MULTI && fabsf(az) <= HIST_R ? 2 : 1;
                for (int kz = 0; kz < nz; kz++)
                    for (int ky = 0; ky < ny; ky++)
                        for (int kx = 0; kx < nx; kx++) {
                            const float xx = kx ? ax : x, yy = ky ? ay : y, zz = kz ? az : z;
                            const float r2 = xx * xx + yy * yy + zz * zz;
                            if (r2 <= r2max) {
                                const int b = min(HIST_BINS - 1, (int)(sqrtf(r2) * binv));
                                const int bi = b + jneg, bj = b + ineg;
                                atomicAdd(&hc[li * 7 + (bi >> 2)], 1u << ((bi & 3) * 8));
                                atomicAdd(&hc[jl * 7 + (bj >> 2)], 1u << ((bj & 3) * 8));
                            }
                        }
            }
            if (st + 1 < ns) { jx = rot1(jx); jy = rot1(jy); jz = rot1(jz); jl = rot1(jl); jneg = rot1(jneg); }
        }
        if (++J == NT) { ++I; J = I; }
    }
    __syncthreads();
    hist_features<T>(a, m, hc);
    __syncthreads();
}

// |F| / Fmax on a log scale, 1 + log10(max(r, 1e-6)) / 6 in [0, 1]: the plain step of an atom is
// proportional to it, so its optimal relative step spans decades.
__device__ __forceinline__ float force_feature(float r) { return 1.f + log10f(fmaxf(r, 1e-6f)) * (1.f / 6.f); }

// Learned step scale of every atom: features = (F / Fmax, log |F| / Fmax, histogram), one MFMA forward
// per 16 atoms per wave.  Returns S = max_i s_i |F_i| (every thread), the normaliser of the proposal.
template <int T>
__device__ __forceinline__ float infer_scales(const BatchArgs& a, const SysMeta& m, const float4* F, float fmax, BlockShared<T>& sh,
                                              const unsigned char* lds) {
    // the MFMA A operands from the copy in LDS (staged at the start of the launch)
    MlpWeights wl = a.mlp;
    wl.w1t = reinterpret_cast<const unsigned short*>(lds + a.wlds);
    wl.w2t = wl.w1t + MLP_H * MLP_IN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    constexpr int W = T / 64;
    const float inv = 1.f / fmax;
    const bf16x4 zero = pack4(0.f, 0.f, 0.f, 0.f);
    float smax = 0.f, slo = 3.0e38f, shi = -3.0e38f;
    // the operands of a group of 16 atoms: force, and this lane's feature words (group g's slice);
    // the next group's loads are issued before the current group's MFMA chain
    auto load = [&](int base, float4& f, bf16x4& x0, bf16x4& x1) {
        const int li = base + r;
        const bool ok = li < m.n;
        const unsigned short* fr = a.feat + (int64_t)(m.atom0 + (ok ? li : 0)) * 32;
        f = ok ? F[m.atom0 + li] : make_float4(0.f, 0.f, 0.f, 0.f);
        x0 = ok ? *reinterpret_cast<const bf16x4*>(fr + 4 * g) : zero;   // g == 0 replaces it below
        x1 = ok ? *reinterpret_cast<const bf16x4*>(fr + 16 + 4 * g) : zero;
    };
    float4 fn;
    bf16x4 x0n, x1n;
    if (wave * 16 < m.n) load(wave * 16, fn, x0n, x1n);
    for (int base = wave * 16; base < m.n; base += W * 16) {
        const int li = base + r;
        const bool ok = li < m.n;
        const float4 f = fn;
        bf16x4 xb[2] = {x0n, x1n};
        if (base + W * 16 < m.n) load(base + W * 16, fn, x0n, x1n);
        const float fl = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z);
        if (g == 0) xb[0] = pack4(f.x * inv, f.y * inv, f.z * inv, force_feature(fl * inv));
        f32x4 z1[4], z2[4];
        bf16x4 h1[4];
        const float sc = mlp_scale(mlp_forward16(wl, xb, z1, h1, z2));
        if (g == 0 && ok) {
            a.scale[m.atom0 + li] = sc;
            smax = fmaxf(smax, sc * fl);
            slo = fminf(slo, sc);
            shi = fmaxf(shi, sc);
        }
    }
    smax = wave_max(smax);
    slo = -wave_max(-slo);
    shi = wave_max(shi);
    if (lane == 0) { sh.fred[wave] = smax; sh.fmin_[wave] = slo; }
    __syncthreads();
    // the scale range, wave by wave (the same lanes reuse fmin_ after the first barrier)
    if (threadIdx.x == 0) {
        float v = 0.f, lo = 3.0e38f;
        for (int w = 0; w < W; w++) { v = fmaxf(v, sh.fred[w]); lo = fminf(lo, sh.fmin_[w]); }
        sh.fres = v;
        sh.res[0] = (double)lo;
    }
    __syncthreads();
    if (lane == 0) sh.fmin_[wave] = shi;
    __syncthreads();
    if (threadIdx.x == 0) {
        float hi = -3.0e38f;
        for (int w = 0; w < W; w++) hi = fmaxf(hi, sh.fmin_[w]);
        sh.uniform = (float)sh.res[0] == hi ? 1 : 0;
    }
    __syncthreads();
    return sh.fres;
}

struct Eval {
    double lj, coul, bond, pme;
    float fmax;
    __device__ double total() const { return lj + coul + bond + pme; }
};

// Forces and energy of the system at positions u (global arrays); writes f and returns the FP64
// energy parts and Fmax (identical in every thread).
template <int T, int UNR, bool MULTI>
__device__ __forceinline__ Eval evaluate(const BatchArgs& a, const SysMeta& m, BlockShared<T>& sh, unsigned char* lds,
                                         const int4* u, float4* f, bool want_hist) {
    const int tid = threadIdx.x;
    const int n = m.n;
    double elj_d = 0.0, ec_d = 0.0, eb_d = 0.0, ep_d = 0.0;
    unsigned long long tk = tid == 0 ? __builtin_amdgcn_s_memrealtime() : 0ull;
    // learned mode, before a feature refresh: the pair pass also counts the neighbour histogram
    // (byte counters in LDS after the waves' j-force slices), copied to hnext afterwards
    const bool hon = __builtin_amdgcn_readfirstlane(want_hist ? 1 : 0) != 0;
    unsigned* hcl = reinterpret_cast<unsigned*>(lds + (T / 64) * 192 * sizeof(float));
    if (hon) {
        for (int k = tid; k < n * 7; k += T) hcl[k] = 0u;
        __syncthreads();
    }
    // ---- cutoff pairs: every unordered pair of 64-atom tiles once (Newton's third law) ----
    // Wave w takes a fixed contiguous range of tile pairs (I <= J).  Lane l holds atom I*64+l;
    // the J tile's atoms and their force accumulators rotate one lane per step (DPP), so at
    // step s lane l meets atom J*64 + (l+s) mod 64.  Each tile pair writes its i- and j-side
    // partial forces; atoms sum them in a fixed tile order, so forces do not depend on the
    // workgroup shape.
    const int NT = (n + 63) >> 6;
    {
        const int lane = tid & 63, wave = tid >> 6;
        constexpr int W = T / 64;
        const int NI = NT * (NT + 1) / 2;
        float* home = reinterpret_cast<float*>(lds) + wave * 192;   // 64 x 3 floats per wave
        float4* part = a.part + m.part0;
        const PairBox bx = {m.Lq[0], m.Lq[1], m.Lq[2], m.hLq[0], m.hLq[1], m.hLq[2], m.L[0], m.L[1], m.L[2],
                            1.f / m.L[0], 1.f / m.L[1], 1.f / m.L[2]};
        const int q0 = (int)((long long)NI * wave / W), q1 = (int)((long long)NI * (wave + 1) / W);
        int I = 0, J = 0;
        {
            int q = q0;
            while (I < NT && q >= NT - I) { q -= NT - I; I++; }
            J = I + q;
        }
        int curI = -1;
        // padding atoms (past n) sit at NaN: every distance test with them is false
        const float pad = __builtin_nanf("");
        float3 wi = make_float3(pad, pad, pad);
        float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
        int ci = 0, bmi = 0;
        bool vi = false;
        HistCtx hh{hcl, 0, 0, 0, 0, hon, false};
        for (int q = q0; q < q1; q++) {
            if (I != curI) {
                curI = I;
                const int li = I * 64 + lane;
                vi = li < n;
                hh.li = li;
                wi = make_float3(pad, pad, pad);
                if (vi) {
                    const int4 uu = u[m.atom0 + li];
                    ci = a.comp[m.atom0 + li];
                    wi = make_float3((float)fmodi(uu.x, bx.Lx) * LIMA_LIMA_TO_NANO_F, (float)fmodi(uu.y, bx.Ly) * LIMA_LIMA_TO_NANO_F,
                                     (float)fmodi(uu.z, bx.Lz) * LIMA_LIMA_TO_NANO_F);
                    pi = a.prm[m.atom0 + li];
                    bmi = a.cbm[m.comp0 + ci];
                }
                hh.ineg = vi && pi.x < 0.f ? HIST_BINS : 0;
            }
            // atom J*64+lane and its accumulator start in this lane
            float jx = pad, jy = pad, jz = pad;
            int jc = -1;
            float jq = 0.f, js = 0.f, je = 0.f;
            {
                const int lj = J * 64 + lane;
                if (lj < n) {
                    const int4 uu = u[m.atom0 + lj];
                    const float4 pj = a.prm[m.atom0 + lj];
                    jx = (float)fmodi(uu.x, bx.Lx) * LIMA_LIMA_TO_NANO_F;
                    jy = (float)fmodi(uu.y, bx.Ly) * LIMA_LIMA_TO_NANO_F;
                    jz = (float)fmodi(uu.z, bx.Lz) * LIMA_LIMA_TO_NANO_F;
                    jc = a.comp[m.atom0 + lj];
                    jq = pj.x; js = pj.y; je = pj.w;
                }
            }
            v3 fi = {0.f, 0.f, 0.f}, fj = {0.f, 0.f, 0.f};
            float elj = 0.f, ec = 0.f;
            const bool diag = I == J;
            const bool near = J - I <= 2;   // compounds hold <= 32 atoms: intra pairs lie within two tiles
            const int ns = diag ? 33 : 64;  // diagonal tile: steps 1..32 (step 32 on half the lanes)
            const float4* tab = sh.ctab;
            if (near) {
#pragma unroll UNR
                for (int st = 0; st < ns; st++) {
                    const int dc = jc - ci;
                    const bool skip = !vi || jc < 0 || (diag && (st == 0 || (st == 32 && lane >= 32)));
                    const bool excl = skip || ((unsigned)(dc + 2) <= 4u && ((bmi >> (dc + 2)) & 1));
                    hh.lj = J * 64 + ((lane + st) & 63);
                    hh.jneg = jq < 0.f ? HIST_BINS : 0;
                    hh.skip = skip;
                    pair_cut<MULTI>(wi, pi, excl, jx, jy, jz, jq, js, je, bx, a, tab, fi, fj, elj, ec, hh);
                    if (st + 1 < ns) {
                        jx = rot1f(jx); jy = rot1f(jy); jz = rot1f(jz); jc = rot1(jc);
                        jq = rot1f(jq); js = rot1f(js); je = rot1f(je);
                        fj.x = rot1f(fj.x); fj.y = rot1f(fj.y); fj.z = rot1f(fj.z);
                    }
                }
            } else {   // tiles more than two apart: no intra or bonded-compound pairs, 64 full steps
                // (padding atoms drop out through their NaN coordinates)
                hh.skip = false;
#pragma unroll UNR
                for (int st = 0; st < 64; st++) {
                    hh.lj = J * 64 + ((lane + st) & 63);
                    hh.jneg = jq < 0.f ? HIST_BINS : 0;
                    pair_cut<MULTI>(wi, pi, false, jx, jy, jz, jq, js, je, bx, a, tab, fi, fj, elj, ec, hh);
                    if (st + 1 < 64) {
                        jx = rot1f(jx); jy = rot1f(jy); jz = rot1f(jz);
                        jq = rot1f(jq); js = rot1f(js); je = rot1f(je);
                        fj.x = rot1f(fj.x); fj.y = rot1f(fj.y); fj.z = rot1f(fj.z);
                    }
                }
            }
            // after ns - 1 rotations lane l holds the accumulator of atom J*64 + (l + ns - 1) mod 64
            const int jidx = (lane + ns - 1) & 63;
            home[jidx] = fj.x; home[64 + jidx] = fj.y; home[128 + jidx] = fj.z;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const float hx = home[lane], hy = home[64 + lane], hz = home[128 + lane];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            part[(q * 2 + 0) * 64 + lane] = make_float4(fi.x, fi.y, fi.z, 0.f);
            part[(q * 2 + 1) * 64 + lane] = make_float4(hx, hy, hz, 0.f);
            elj_d += (double)elj;
            ec_d += (double)ec;
            if (++J == NT) { ++I; J = I; }
        }
    }
    __syncthreads();
    if (hon)   // before the PME pass reuses the LDS
        for (int k = tid; k < n * 7; k += T) a.hnext[(int64_t)m.atom0 * 7 + k] = hcl[k];
    if (tid == 0) { const unsigned long long t = __builtin_amdgcn_s_memrealtime(); sh.ph[0] += t - tk; tk = t; }
    // ---- per atom: tile-pair partials in a fixed order, intra pairs, bonded terms ----
    {
        const float4* part = a.part + m.part0;
        for (int li = tid; li < n; li += T) {
            const int A = li >> 6, la = li & 63;
            v3 acc = {0.f, 0.f, 0.f};
            for (int I2 = 0; I2 <= A; I2++) {   // j side of (I2, A)
                const int q = I2 * NT - I2 * (I2 - 1) / 2 + (A - I2);
                const float4 v = part[(q * 2 + 1) * 64 + la];
                acc.x += v.x; acc.y += v.y; acc.z += v.z;
            }
            for (int J2 = A; J2 < NT; J2++) {   // i side of (A, J2)
                const int q = A * NT - A * (A - 1) / 2 + (J2 - A);
                const float4 v = part[(q * 2 + 0) * 64 + la];
                acc.x += v.x; acc.y += v.y; acc.z += v.z;
            }
            const int gi = m.atom0 + li;
            const int ci = a.comp[gi];
            float elj = 0.f, ec = 0.f, eb = 0.f;
            intra_atom(a, m, u, li, ci, a.cbm[m.comp0 + ci], a.prm[gi], acc, elj, ec);
            bonded_atom(a, m, u, gi, acc, eb);
            f[gi] = make_float4(acc.x, acc.y, acc.z, 0.f);
            elj_d += (double)elj;
            ec_d += (double)ec;
            eb_d += (double)eb;
        }
    }
    // ---- PME ----
    __syncthreads();
    if (tid == 0) { const unsigned long long t = __builtin_amdgcn_s_memrealtime(); sh.ph[1] += t - tk; tk = t; }
    if (m.elec) {
        int* qg = reinterpret_cast<int*>(lds);
        float* phi = reinterpret_cast<float*>(lds);
        const int K0 = m.K[0], K1 = m.K[1], K2 = m.K[2], H = m.H;
        for (int g = tid; g < m.NG; g += T) qg[g] = 0;
        for (int d = 0; d < 3; d++)
            for (int k = tid; k < m.K[d]; k += T) sh.tw[d][k] = a.twid[m.tw0[d] + k];
        __syncthreads();
        for (int li = tid; li < n; li += T) {   // spread q w_x w_y w_z (int32 fixed point)
            const float q = a.prm[m.atom0 + li].x;
            if (q == 0.f) continue;
            const int4 uu = u[m.atom0 + li];
            int ix[3];
            float fr[3], wx[4], wy[4], wz[4];
            grid_coord((uint32_t)fmodi(uu.x, m.Lq[0]), m.Lq[0], K0, &ix[0], &fr[0]);
            grid_coord((uint32_t)fmodi(uu.y, m.Lq[1]), m.Lq[1], K1, &ix[1], &fr[1]);
            grid_coord((uint32_t)fmodi(uu.z, m.Lq[2]), m.Lq[2], K2, &ix[2], &fr[2]);
            bspline(fr[0], wx); bspline(fr[1], wy); bspline(fr[2], wz);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int Z = wrapi(ix[2] - 1 + c, K2);
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int Y = wrapi(ix[1] - 1 + b, K1);
                    const float qyz = q * wy[b] * wz[c];
                    int* row = qg + (Z * K1 + Y) * K0;
#pragma unroll
                    for (int x = 0; x < 4; x++) atomicAdd(&row[wrapi(ix[0] - 1 + x, K0)], __float2int_rn(qyz * wx[x] * FIX_SCALE));
                }
            }
        }
        __syncthreads();
        // The transforms compute one output coefficient per thread (a length-K sum with the
        // twiddles from LDS), so no thread holds a whole line, and every pass has H K K' work
        // items for the workgroup.  The half spectrum ping-pongs between spec and spec2.
        float2* spec = a.spec + m.spec0;
        float2* spec2 = a.spec2 + m.spec0;
        const float rH = 1.f / (float)H, rK0 = 1.f / (float)K0;
        // item it = row * n + col, split with a float reciprocal (exact below 2^22 items)
        auto split = [](int it, float rn, int n, int* col) { const int r = (int)(((float)it + 0.5f) * rn); *col = it - r * n; return r; };
        // x forward, real -> half spectrum: out (row l = (z, y), k) = sum_j rho(l, j) W0^(j k)
        for (int it = tid; it < K1 * K2 * H; it += T) {
            int k;
            const int l = split(it, rH, H, &k);
            const int* src = qg + l * K0;
            float re = 0.f, im = 0.f;
            int idx = 0;
            for (int j = 0; j < K0; j++) {
                const float v = (float)src[j] * m.fix_to_rho;
                const float2 w = sh.tw[0][idx];
                re = fmaf(v, w.x, re);
                im = fmaf(v, w.y, im);
                idx += k;
                idx -= idx >= K0 ? K0 : 0;
            }
            spec[l * H + k] = make_float2(re, im);
        }
        __syncthreads();
        // one complex pass of length K along y (stride H, outer z with stride K1 H) or z (stride
        // K1 H, outer y with stride H): item (outer o, k, x), out[o ostride + x + k stride] =
        // sum_j in[o ostride + x + j stride] W^(j k)
        auto pass = [&](const float2* in, float2* out, const float2* tw, int K, int stride, int nout, int ostride, bool inv,
                        bool green) {
            const float rK = 1.f / (float)K;
            for (int it = tid; it < H * K * nout; it += T) {
                int x, k;
                const int r = split(it, rH, H, &x);
                const int o = split(r, rK, K, &k);
                const int line0 = o * ostride + x;
                float re = 0.f, im = 0.f;
                int idx = 0;
                for (int j = 0; j < K; j++) {
                    const float2 v = in[line0 + j * stride];
                    const float2 w = tw[idx];
                    const float wi = inv ? -w.y : w.y;
                    re = fmaf(v.x, w.x, fmaf(-v.y, wi, re));
                    im = fmaf(v.x, wi, fmaf(v.y, w.x, im));
                    idx += k;
                    idx -= idx >= K ? K : 0;
                }
                const int oi = line0 + k * stride;
                if (green) { const float gk = a.green[m.spec0 + oi]; re *= gk; im *= gk; }
                out[oi] = make_float2(re, im);
            }
            __syncthreads();
        };
        pass(spec, spec2, sh.tw[1], K1, H, K2, K1 * H, false, false);    // y forward
        pass(spec2, spec, sh.tw[2], K2, K1 * H, K1, H, false, true);     // z forward times the influence function
        pass(spec, spec2, sh.tw[2], K2, K1 * H, K1, H, true, false);     // z inverse
        pass(spec2, spec, sh.tw[1], K1, H, K2, K1 * H, true, false);     // y inverse
        // x inverse, half spectrum -> real potential (Hermitian sum) / NG into LDS
        {
            const bool even = (K0 & 1) == 0;
            for (int it = tid; it < K1 * K2 * K0; it += T) {
                int xj;
                const int l = split(it, rK0, K0, &xj);
                const float2* src = spec + l * H;
                float s = src[0].x;
                int idx = xj;
                for (int k = 1; k < H; k++) {
                    const float2 v = src[k];
                    const float2 w = sh.tw[0][idx];   // exp(-i ...); the inverse uses the conjugate
                    const float re = v.x * w.x + v.y * w.y;
                    s += (even && k == H - 1) ? re : 2.f * re;
                    idx += xj;
                    idx -= idx >= K0 ? K0 : 0;
                }
                phi[l * K0 + xj] = s * m.inv_ng;
            }
        }
        __syncthreads();
        // gather: phi and the finite-difference field, force q E, energy q phi / 2
        for (int li = tid; li < n; li += T) {
            const float q = a.prm[m.atom0 + li].x;
            if (q == 0.f) continue;
            const int4 uu = u[m.atom0 + li];
            int ix[3];
            float fr[3], wx[4], wy[4], wz[4];
            grid_coord((uint32_t)fmodi(uu.x, m.Lq[0]), m.Lq[0], K0, &ix[0], &fr[0]);
            grid_coord((uint32_t)fmodi(uu.y, m.Lq[1]), m.Lq[1], K1, &ix[1], &fr[1]);
            grid_coord((uint32_t)fmodi(uu.z, m.Lq[2]), m.Lq[2], K2, &ix[2], &fr[2]);
            bspline(fr[0], wx); bspline(fr[1], wy); bspline(fr[2], wz);
            float w6x[6], w6y[6], w6z[6], d6x[6], d6y[6], d6z[6];
            gather_weights(wx, w6x, d6x);
            gather_weights(wy, w6y, d6y);
            gather_weights(wz, w6z, d6z);
            // per row: the x sums with the weights and their differences (two FMAs per node), then
            // the row's y and z factors; the z planes in a loop (weights picked, not indexed)
            int xs[6];
#pragma unroll
            for (int mx = 0; mx < 6; mx++) xs[mx] = wrapi(wrapi(ix[0] - 2 + mx, K0), K0);
            float ph = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
#pragma unroll 1
            for (int mz = 0; mz < 6; mz++) {
                const int Z = wrapi(wrapi(ix[2] - 2 + mz, K2), K2);
                float wz = 0.f, dz = 0.f;
#pragma unroll
                for (int c = 0; c < 6; c++) {
                    wz = mz == c ? w6z[c] : wz;
                    dz = mz == c ? d6z[c] : dz;
                }
#pragma unroll
                for (int my = 0; my < 6; my++) {
                    const int Y = wrapi(wrapi(ix[1] - 2 + my, K1), K1);
                    const float* row = phi + (Z * K1 + Y) * K0;
                    float rw = 0.f, rd = 0.f;
#pragma unroll
                    for (int mx = 0; mx < 6; mx++) {
                        const float v = row[xs[mx]];
                        rw = fmaf(w6x[mx], v, rw);
                        rd = fmaf(d6x[mx], v, rd);
                    }
                    const float wyz = w6y[my] * wz, dyz = d6y[my] * wz, wdz = w6y[my] * dz;
                    ph = fmaf(wyz, rw, ph);
                    ex = fmaf(wyz, rd, ex);
                    ey = fmaf(dyz, rw, ey);
                    ez = fmaf(wdz, rw, ez);
                }
            }
            float4 fo = f[m.atom0 + li];
            fo.x += q * ex * m.inv2h[0];
            fo.y += q * ey * m.inv2h[1];
            fo.z += q * ez * m.inv2h[2];
            f[m.atom0 + li] = fo;
            ep_d += (double)(0.5f * q * ph);
        }
    }
    // ---- Fmax and the FP64 energy sums ----
    __syncthreads();
    if (tid == 0) { const unsigned long long t = __builtin_amdgcn_s_memrealtime(); sh.ph[2] += t - tk; tk = t; }
    float fm = 0.f;
    for (int li = tid; li < n; li += T) {
        const float4 fo = f[m.atom0 + li];
        fm = fmaxf(fm, sqrtf(fo.x * fo.x + fo.y * fo.y + fo.z * fo.z));
    }
    double v[4] = {elj_d, ec_d, eb_d, ep_d};
    block_reduce<T>(sh, v, fm);
    if (tid == 0) sh.ph[3] += __builtin_amdgcn_s_memrealtime() - tk;
    return Eval{v[0], v[1], v[2], v[3], sh.fres};
}

// Write the sampled atoms of the accepted state: features (F / Fmax, |F| / Fmax, histogram), the
// plain step h F / Fmax the minimiser would take from here, and the Coord.
template <int T>
__device__ __forceinline__ void record_state(const BatchArgs& a, const SysMeta& m, int sys, int state, const int4* u, const float4* F, float fmax,
                             double h) {
    const int2 ri = a.rec_info[sys];
    const float inv = 1.f / fmax, k = (float)(h / (double)fmax);
    for (int q = threadIdx.x; q < ri.y; q += T) {
        const int li = q * a.rec_stride;
        float* e = a.rec + ri.x + ((int64_t)state * ri.y + q) * REC_W;
        const float4 f = F[m.atom0 + li];
        const int4 x = u[m.atom0 + li];
        e[0] = f.x * inv; e[1] = f.y * inv; e[2] = f.z * inv; e[3] = force_feature(sqrtf(f.x * f.x + f.y * f.y + f.z * f.z) * inv);
        const unsigned short* fr = a.feat + (int64_t)(m.atom0 + li) * 32;
        for (int b = 4; b < 32; b++) e[b] = bf2f(fr[b]);
        e[32] = k * f.x; e[33] = k * f.y; e[34] = k * f.z;
        e[35] = __int_as_float(x.x); e[36] = __int_as_float(x.y); e[37] = __int_as_float(x.z);
        e[38] = f.x; e[39] = f.y; e[40] = f.z;
        for (int b = 41; b < REC_W; b++) e[b] = 0.f;
    }
}

// Trial step from the accepted state: x' = x + trunc(h s F / norm), s = 1 (plain, norm = Fmax) or the
// learned scale (norm = max s |F|).
template <int T>
__device__ __forceinline__ void trial_step(const SysMeta& m, const int4* uc, const float4* fc, int4* ut, const float* scale, double h,
                                           float fmax) {
    const float k = (float)(h / (double)fmax);
    for (int li = threadIdx.x; li < m.n; li += T) {
        const int4 x = uc[li];
        const float4 g = fc[li];
        const float kk = scale ? k * scale[li] : k;
        ut[li] = make_int4(x.x + (int)fminf(fmaxf(kk * g.x * LIMA_NANO_TO_LIMA_F, -2.0e8f), 2.0e8f),
                           x.y + (int)fminf(fmaxf(kk * g.y * LIMA_NANO_TO_LIMA_F, -2.0e8f), 2.0e8f),
                           x.z + (int)fminf(fmaxf(kk * g.z * LIMA_NANO_TO_LIMA_F, -2.0e8f), 2.0e8f), 0);
    }
}

template <int T, int UNR, bool MULTI>
__device__ __forceinline__ void minimize_system(const BatchArgs& a, const SysMeta& m, int sys, SysOut* out, BlockShared<T>& sh,
                                                unsigned char* lds) {
    const int tid = threadIdx.x;
    if (tid == 0) for (int q = 0; q < 7; q++) sh.ph[q] = 0ull;
    const bool learned = a.mode == MODE_LEARNED, record = a.mode == MODE_RECORD;
    Eval acc_e{0.0, 0.0, 0.0, 0.0, 0.f};
    double h = a.h0;
    int cur = 0, it = 0, nacc = 0, nevals = 0, nlearn = 0, nrec = 0;
    // pass kinds: 0 the starting state, 1 a learned proposal, 2 the plain step
    int kind = 0;
    float snorm = 1.f;
    bool conv = false;
    bool hist_ok = false;   // hnext holds the neighbour histogram of the current accepted state
    for (;;) {
        const int tgt = kind == 0 ? 0 : 1 - cur;
        int4* ut = tgt ? a.u1 : a.u0;
        float4* ft = tgt ? a.f1 : a.f0;
        if (kind != 0) {
            const int4* uc = (cur ? a.u1 : a.u0) + m.atom0;
            const float4* fc = (cur ? a.f1 : a.f0) + m.atom0;
            if (kind == 1) {
                unsigned long long t0 = tid == 0 ? __builtin_amdgcn_s_memrealtime() : 0ull;
                if (it % a.feat_interval == 0) {
                    // the histogram counted in the pair pass of the accepted state's evaluation, or
                    // (when that evaluation counted none) a pass of its own
                    if (hist_ok) hist_features<T>(a, m, a.hnext + (int64_t)m.atom0 * 7);
                    else hist_pass<T, MULTI>(a, m, lds, cur ? a.u1 : a.u0);
                    __syncthreads();
                    if (tid == 0) { const unsigned long long t1 = __builtin_amdgcn_s_memrealtime(); sh.ph[5] += t1 - t0; t0 = t1; }
                }
                snorm = infer_scales<T>(a, m, cur ? a.f1 : a.f0, acc_e.fmax, sh, lds);
                if (tid == 0) sh.ph[6] += __builtin_amdgcn_s_memrealtime() - t0;
                // every atom got the same scale: the proposal is the plain step's direction with
                // the same largest displacement, so the plain step is taken (one evaluation, and a
                // refusal needs no second one)
                if (sh.uniform) kind = 2;
            }
            const unsigned long long ts = tid == 0 ? __builtin_amdgcn_s_memrealtime() : 0ull;
            // plain: h F / Fmax; learned: h s F / max(s |F|), the same largest displacement h
            trial_step<T>(m, uc, fc, ut + m.atom0, kind == 1 ? a.scale + m.atom0 : nullptr, h, kind == 1 ? snorm : acc_e.fmax);
            __syncthreads();
            if (tid == 0) sh.ph[4] += __builtin_amdgcn_s_memrealtime() - ts;
        }
        // the evaluations that may produce the state of the next feature refresh count its histogram
        const bool want_hist = learned && (kind == 0 || (it + 1) % a.feat_interval == 0);
        const Eval ev = evaluate<T, UNR, MULTI>(a, m, sh, lds, ut, ft, want_hist);
        nevals++;
        bool step_done = false, accepted = false;
        if (kind == 0) {
            acc_e = ev;
            accepted = true;   // for recording the starting state
        } else if (ev.total() < acc_e.total()) {   // FP64 acceptance
            cur = tgt;
            acc_e = ev;
            accepted = true;
            step_done = true;
            nlearn += kind == 1;
        } else if (kind == 2) {
            step_done = true;
        }
        if (want_hist) hist_ok = accepted;   // hnext now holds this trial's histogram
        else if (accepted) hist_ok = false;  // the current state changed without one
        if (kind != 0 && step_done) {
            if (accepted) {
                if (kind == 2 || a.grow_on_learned) h *= 1.2;
                nacc++;
            } else {
                h *= 0.2;
            }
            it++;
        }
        if (record && accepted && nrec < a.rec_cap) {
            hist_pass<T, MULTI>(a, m, lds, cur ? a.u1 : a.u0);
            record_state<T>(a, m, sys, nrec++, cur ? a.u1 : a.u0, cur ? a.f1 : a.f0, acc_e.fmax, h);
        }
        if (kind == 1 && !step_done) { kind = 2; continue; }   // proposal refused: plain step, same iteration
        conv = (double)acc_e.fmax < a.tol;
        if (conv || it >= a.max_iter || h < 1e-12) break;
        kind = learned ? 1 : 2;
    }
    if (cur == 1) {   // the accepted state goes back to buffer 0
        __syncthreads();
        for (int li = tid; li < m.n; li += T) {
            a.u0[m.atom0 + li] = a.u1[m.atom0 + li];
            a.f0[m.atom0 + li] = a.f1[m.atom0 + li];
        }
    }
    if (tid == 0) {
        out->e_acc = acc_e.total() + m.eself;
        out->e_parts[0] = acc_e.lj;
        out->e_parts[1] = acc_e.coul;
        out->e_parts[2] = acc_e.bond;
        out->e_parts[3] = acc_e.pme;
        out->fmax = acc_e.fmax;
        out->h = (float)h;
        out->iterations = it;
        out->accepted = nacc;
        out->converged = conv ? 1 : 0;
        out->cur = 0;
        out->evaluations = nevals;
        out->learned_accepted = nlearn;
        out->rec_iters = nrec;
        for (int q = 0; q < 7; q++) out->ph[q] = sh.ph[q];
    }
}

template <int T, int UNR>
__global__ __launch_bounds__(T) void k_em_batch(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ BlockShared<T> sh;
    const int s = a.order[blockIdx.x];
    const SysMeta m = a.meta[s];
    SysOut* out = a.out + s;
    if (threadIdx.x == 0) out->t0 = __builtin_amdgcn_s_memrealtime();
    for (int i = threadIdx.x; i < 2 * (CTAB_N + 1); i += T) sh.ctab[i] = a.ctab[i];
    if (a.mode == MODE_LEARNED) {   // model weights, 8 bytes per thread and step
        uint2* wd = reinterpret_cast<uint2*>(lds + a.wlds);
        const uint2* w1 = reinterpret_cast<const uint2*>(a.mlp.w1t);
        const uint2* w2 = reinterpret_cast<const uint2*>(a.mlp.w2t);
        constexpr int n1 = MLP_H * MLP_IN / 4, n2 = MLP_H * MLP_H / 4;
        for (int i = threadIdx.x; i < n1 + n2; i += T) wd[i] = i < n1 ? w1[i] : w2[i - n1];
    }
    __syncthreads();
    if (m.multi) minimize_system<T, UNR, true>(a, m, s, out, sh, lds);
    else minimize_system<T, UNR, false>(a, m, s, out, sh, lds);
    __syncthreads();
    if (threadIdx.x == 0) out->t1 = __builtin_amdgcn_s_memrealtime();
}

// ---------------------------------------------------------------------------------------------

struct Shape { int threads, unroll; };
constexpr Shape SHAPES[] = {{256, 1}, {256, 2}, {256, 4}, {512, 1}, {512, 2}, {512, 4}, {1024, 1}, {1024, 2}, {1024, 4}};
constexpr int NSHAPES = (int)(sizeof(SHAPES) / sizeof(SHAPES[0]));

typedef void (*KernelFn)(BatchArgs);
KernelFn shape_kernel(int i) {
    switch (i) {
        case 0: return k_em_batch<256, 1>;
        case 1: return k_em_batch<256, 2>;
        case 2: return k_em_batch<256, 4>;
        case 3: return k_em_batch<512, 1>;
        case 4: return k_em_batch<512, 2>;
        case 5: return k_em_batch<512, 4>;
        case 6: return k_em_batch<1024, 1>;
        case 7: return k_em_batch<1024, 2>;
        default: return k_em_batch<1024, 4>;
    }
}

int smooth235(int n) { for (int p : {2, 3, 5}) while (n % p == 0) n /= p; return n == 1; }
int derive_grid_b(double L, double spacing) {
    int n = (int)std::ceil(L / spacing - 1e-9);
    if (n < 4) n = 4;
    while (!smooth235(n)) n++;
    return n;
}
int64_t fdiv64(int64_t a, int64_t b) { int64_t q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) --q; return q; }

template <class T>
T* dcopy(const std::vector<T>& v) {
    T* p = nullptr;
    if (v.empty()) return nullptr;
    HIP_CHECK(hipMalloc(&p, v.size() * sizeof(T)));
    HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

}  // namespace

struct Batch {
    int nsys = 0;
    int64_t N = 0;
    lima_simparams prm{};
    PairConst pc{};
    float rc = 1.2f;
    std::vector<SysMeta> meta;
    std::vector<int> h_ncomp;                  // compounds per system
    std::vector<std::vector<int>> h_cn;        // per system: atoms per compound
    std::vector<std::vector<int>> h_key;       // per system: key particle per compound
    std::vector<int> h_ntiny;
    std::vector<int4> u_init;
    // device
    SysMeta* d_meta = nullptr;
    int* d_order = nullptr;
    int4 *u0 = nullptr, *u1 = nullptr, *u_start = nullptr;
    float4 *f0 = nullptr, *f1 = nullptr;
    float4* prm4 = nullptr;
    int* comp = nullptr;
    int4* cinfo = nullptr;
    int* cbm = nullptr;
    uint32_t* clut = nullptr;
    int *inc_off = nullptr, *inc = nullptr;
    int2* bid = nullptr; float2* bp = nullptr;
    int4* aid = nullptr; float4* ap = nullptr;
    int4* did = nullptr; float4* dp = nullptr;
    int4* iid = nullptr; float2* ip = nullptr;
    int2* pid = nullptr; float2* pp = nullptr;
    float* green = nullptr;
    float2* spec = nullptr;
    float2* spec2 = nullptr;
    float4* part = nullptr;
    float2* twid = nullptr;
    float4* ctab = nullptr;
    SysOut* out = nullptr;
    unsigned short* feat = nullptr;   // [N][32] BF16 features (learned and recording modes)
    unsigned* hnext = nullptr;        // [N][7] histogram bytes counted in a pair pass (learned mode)
    float* scale = nullptr;           // [N]
    float* rec = nullptr;             // recorded states
    int2* rec_info = nullptr;         // per system (offset, sampled atoms)
    int64_t rec_floats = 0;
    int rec_stride = 0, rec_cap = 0;
    std::vector<int2> h_rec_info;
    hipStream_t streams[NSHAPES] = {};
    lima_batch_tuning tuning{};
    bool tuned = false;
    bool forces_valid = false;
};

namespace {

void batch_free(Batch& b) {
    for (auto& s : b.streams) if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); s = nullptr; }
    void* ptrs[] = {b.d_meta, b.d_order, b.u0, b.u1, b.u_start, b.f0, b.f1, b.prm4, b.comp, b.cinfo, b.cbm, b.clut, b.inc_off, b.inc,
                    b.bid, b.bp, b.aid, b.ap, b.did, b.dp, b.iid, b.ip, b.pid, b.pp, b.green, b.spec, b.spec2, b.part, b.twid, b.ctab, b.out,
                    b.feat, b.hnext, b.scale, b.rec, b.rec_info};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}

void batch_create(Batch& B, const lima_box_desc* boxes, int nsys, const lima_simparams* p) {
    if (!boxes || !p || nsys <= 0) throw LimaFail(LIMA_ERR_BAD_INPUT, "null descriptor or empty batch");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw LimaFail(LIMA_ERR_NO_DEVICE, "no HIP device visible");
    if (p->bc_select != LIMA_BC_PBC) throw LimaFail(LIMA_ERR_UNSUPPORTED, "only periodic boundary conditions are implemented");
    if (!(p->cutoff_nm > 0.f)) throw LimaFail(LIMA_ERR_BAD_INPUT, "cutoff must be positive");
    B.prm = *p;
    B.nsys = nsys;
    B.rc = p->cutoff_nm;
    const bool elec = p->enable_electrostatics != 0;
    B.pc.kc = (float)(8.9875517873681764e9 / 1e-9 / 6.02214076e23 * 1000.f * 1000.f);   // PhysicsUtils.cuh:5
    B.pc.kappa = 3.f / B.rc;                                                             // PhysicsUtils.cuh:60-62
    B.pc.two_kappa_sqrtpi = 2.f * B.pc.kappa / sqrtf(LIMA_PI_F);
    {   // erfc(kappa r) and H(r) with their r-derivatives, in double (the explicit term uses LIMA's
        // PI_L as the pair force does; the derivative of erfc is exact); cubic Hermite per interval
        const double kap = (double)B.pc.kappa, cc = (double)B.pc.two_kappa_sqrtpi, hr = (double)B.rc / CTAB_N;
        const double de = 2.0 * kap / std::sqrt(M_PI);
        auto A = [&](double r) { return std::erfc(kap * r); };
        auto dA = [&](double r) { return -de * std::exp(-kap * kap * r * r); };
        auto H = [&](double r) { return std::erfc(kap * r) + cc * r * std::exp(-kap * kap * r * r); };
        auto dH = [&](double r) { return (cc * (1.0 - 2.0 * kap * kap * r * r) - de) * std::exp(-kap * kap * r * r); };
        auto herm = [&](double f0, double f1, double d0, double d1) {
            d0 *= hr; d1 *= hr;
            return make_float4((float)f0, (float)d0, (float)(3.0 * (f1 - f0) - 2.0 * d0 - d1), (float)(2.0 * (f0 - f1) + d0 + d1));
        };
        std::vector<float4> tab(2 * (CTAB_N + 1));
        for (int i = 0; i < CTAB_N; i++) {
            const double r0 = i * hr, r1 = (i + 1) * hr;
            tab[2 * i] = herm(A(r0), A(r1), dA(r0), dA(r1));
            tab[2 * i + 1] = herm(H(r0), H(r1), dH(r0), dH(r1));
        }
        tab[2 * CTAB_N] = make_float4((float)A(B.rc), 0.f, 0.f, 0.f);   // r == cutoff exactly
        tab[2 * CTAB_N + 1] = make_float4((float)H(B.rc), 0.f, 0.f, 0.f);
        B.ctab = dcopy(tab);
    }
    const double spacing = p->pme_spacing_nm > 0 ? p->pme_spacing_nm : 0.12;

    std::vector<int4> u;
    std::vector<float4> prm;
    std::vector<int> comp, cbm;
    std::vector<int4> cinfo;
    std::vector<uint32_t> clut;
    std::vector<int2> bid, pid; std::vector<float2> bp, pp, ip;
    std::vector<int4> aid, did, iid; std::vector<float4> ap, dp;
    std::vector<std::vector<int>> incl;    // per global atom
    std::vector<float> green;
    std::vector<float2> twid;
    int spec_total = 0;
    int64_t part_total = 0;
    B.meta.resize(nsys);
    B.h_ncomp.resize(nsys);
    B.h_cn.resize(nsys);
    B.h_key.resize(nsys);
    B.h_ntiny.resize(nsys);
    for (int s = 0; s < nsys; s++) {
        const lima_box_desc* b = &boxes[s];
        SysMeta& m = B.meta[s];
        std::memset(&m, 0, sizeof(m));
        if (b->n_compounds < 0 || b->n_tinymols < 0) throw LimaFail(LIMA_ERR_BAD_INPUT, "negative counts");
        const int nC = b->n_compounds, nT = b->n_tinymols;
        m.atom0 = (int)u.size();
        m.comp0 = (int)cinfo.size();
        m.nc = nC + nT;     // tinymols become one-atom compounds
        B.h_ncomp[s] = nC;
        B.h_ntiny[s] = nT;
        double L[3];
        for (int d = 0; d < 3; d++) {
            const long long Lq = llround(b->box_nm[d] * 1e8);
            if (Lq <= 0 || Lq >= 1000000000ll) throw LimaFail(LIMA_ERR_BAD_INPUT, "batched systems need box edges in (0, 10) nm");
            m.Lq[d] = (int)Lq;
            m.hLq[d] = (int)(Lq / 2);
            L[d] = (double)Lq / 1e8;
            m.L[d] = (float)L[d];
            if (L[d] <= (double)B.rc) throw LimaFail(LIMA_ERR_BAD_INPUT, "box edge must exceed the cutoff");
            if (L[d] < 2.0 * (double)B.rc) m.multi = 1;
        }
        std::vector<int> first(nC + 1, 0);
        for (int c = 0; c < nC; c++) {
            const int n = b->compound_n_particles[c];
            if (n < 0 || n > 32) throw LimaFail(LIMA_ERR_BAD_INPUT, "compound particle count out of [0, 32]");
            first[c + 1] = first[c] + n;
        }
        m.n = first[nC] + nT;
        if (m.n <= 0 || m.n > MAX_ATOMS) throw LimaFail(LIMA_ERR_BAD_INPUT, "batched systems hold 1 .. 8192 atoms");
        B.h_cn[s].assign(b->compound_n_particles, b->compound_n_particles + nC);
        B.h_key[s].resize(nC);
        double qsq = 0.0;
        int ncharged = 0;
        std::vector<float> qv;
        for (int c = 0; c < nC; c++) {
            B.h_key[s][c] = b->compound_key_particle ? b->compound_key_particle[c] : 0;
            for (int k = 0; k < b->compound_n_particles[c]; k++) {
                const size_t ps = (size_t)c * 32 + k;
                int64_t uu[3];
                for (int d = 0; d < 3; d++) uu[d] = (int64_t)b->compound_origo[c * 3 + d] * 100000000ll + b->particle_coord[ps * 3 + d];
                for (int d = 0; d < 3; d++)
                    if (uu[d] < -1000000000ll || uu[d] > 1000000000ll) throw LimaFail(LIMA_ERR_BAD_INPUT, "batched positions must lie within 10 nm of the origin");
                u.push_back(make_int4((int)uu[0], (int)uu[1], (int)uu[2], 0));
                const float q = elec ? b->particle_charge[ps] : 0.f;
                const float sg = b->particle_sigma[ps], ep = b->particle_epsilon[ps];
                if (sg < 0.f || ep < 0.f) throw LimaFail(LIMA_ERR_BAD_INPUT, "negative LJ parameter");
                prm.push_back(make_float4(q, 0.5f * sg, sqrtf(ep), sqrtf(24.f * ep)));
                comp.push_back(c);
                qsq += (double)q * q;
                ncharged += q != 0.f;
            }
        }
        for (int t = 0; t < nT; t++) {
            int64_t uu[3];
            for (int d = 0; d < 3; d++) uu[d] = (int64_t)b->tinymol_block[t * 3 + d] * 100000000ll + b->tinymol_coord[t * 3 + d];
            u.push_back(make_int4((int)uu[0], (int)uu[1], (int)uu[2], 0));
            const float sg = b->tinymol_sigma[t], ep = b->tinymol_epsilon[t];
            prm.push_back(make_float4(0.f, 0.5f * sg, sqrtf(ep), sqrtf(24.f * ep)));
            comp.push_back(nC + t);
        }
        // compounds: (first, count), bonded mask, exclusion rows (self bits on the diagonal block)
        const uint32_t* lut = b->bonded_lut;
        auto lutrow = [&](int c, int k, int i) -> uint32_t { return lut ? lut[((size_t)c * 5 + k) * 32 + i] : (k == 2 ? (1u << i) : 0u); };
        for (int c = 0; c < nC + nT; c++) {
            const bool tiny = c >= nC;
            const int cf = tiny ? first[nC] + (c - nC) : first[c];
            const int ck = cf + (tiny || !b->compound_key_particle ? 0 : b->compound_key_particle[c]);
            cinfo.push_back(make_int4(cf, tiny ? 1 : b->compound_n_particles[c], ck, 0));
            int bm = 1 << 2;
            for (int k = 0; k < 5; k++) {
                const int Bc = c + k - 2;
                for (int i = 0; i < 32; i++) {
                    uint32_t r = 0;
                    if (!tiny && Bc >= 0 && Bc < nC) r = lutrow(c, k, i);
                    if (k == 2) r |= 1u << i;
                    clut.push_back(r);
                    if (k != 2 && r) bm |= 1 << k;
                }
            }
            cbm.push_back(bm);
        }
        // bonded terms: ids compound*32 + slot -> global atom index; incidences per atom
        incl.resize(u.size());
        auto gid = [&](int32_t id) -> int {
            const int c = id / 32, k = id % 32;
            if (id < 0 || c >= nC || k >= b->compound_n_particles[c]) throw LimaFail(LIMA_ERR_BAD_INPUT, "bonded particle id out of range");
            return m.atom0 + first[c] + k;
        };
        auto add_inc = [&](int atom, int type, int role, int term) {
            if (term >= (1 << 27)) throw LimaFail(LIMA_ERR_BAD_INPUT, "too many bonded terms in the batch");
            incl[atom].push_back((type << 29) | (role << 27) | term);
        };
        for (int k = 0; k < b->n_bonds; k++) {
            const int t = (int)bid.size();
            bid.push_back(make_int2(gid(b->bond_ids[2 * k]), gid(b->bond_ids[2 * k + 1])));
            bp.push_back(make_float2(b->bond_params[2 * k], b->bond_params[2 * k + 1]));
            add_inc(bid[t].x, T_BOND, 0, t); add_inc(bid[t].y, T_BOND, 1, t);
        }
        for (int k = 0; k < b->n_angles; k++) {
            const int t = (int)aid.size();
            aid.push_back(make_int4(gid(b->angle_ids[3 * k]), gid(b->angle_ids[3 * k + 1]), gid(b->angle_ids[3 * k + 2]), 0));
            ap.push_back(make_float4(b->angle_params[4 * k], b->angle_params[4 * k + 1], b->angle_params[4 * k + 2], b->angle_params[4 * k + 3]));
            add_inc(aid[t].x, T_ANGLE, 0, t); add_inc(aid[t].y, T_ANGLE, 1, t); add_inc(aid[t].z, T_ANGLE, 2, t);
        }
        for (int k = 0; k < b->n_dihedrals; k++) {
            const int t = (int)did.size();
            did.push_back(make_int4(gid(b->dihedral_ids[4 * k]), gid(b->dihedral_ids[4 * k + 1]), gid(b->dihedral_ids[4 * k + 2]),
                                    gid(b->dihedral_ids[4 * k + 3])));
            dp.push_back(make_float4(b->dihedral_params[3 * k], b->dihedral_params[3 * k + 1], b->dihedral_params[3 * k + 2], 0.f));
            add_inc(did[t].x, T_DIH, 0, t); add_inc(did[t].y, T_DIH, 1, t); add_inc(did[t].z, T_DIH, 2, t); add_inc(did[t].w, T_DIH, 3, t);
        }
        for (int k = 0; k < b->n_impropers; k++) {
            const int t = (int)iid.size();
            iid.push_back(make_int4(gid(b->improper_ids[4 * k]), gid(b->improper_ids[4 * k + 1]), gid(b->improper_ids[4 * k + 2]),
                                    gid(b->improper_ids[4 * k + 3])));
            ip.push_back(make_float2(b->improper_params[2 * k], b->improper_params[2 * k + 1]));
            add_inc(iid[t].x, T_IMP, 0, t); add_inc(iid[t].y, T_IMP, 1, t); add_inc(iid[t].z, T_IMP, 2, t); add_inc(iid[t].w, T_IMP, 3, t);
        }
        for (int k = 0; k < b->n_pairs; k++) {
            const int t = (int)pid.size();
            pid.push_back(make_int2(gid(b->pair_ids[2 * k]), gid(b->pair_ids[2 * k + 1])));
            pp.push_back(make_float2(b->pair_params[2 * k], b->pair_params[2 * k + 1]));
            add_inc(pid[t].x, T_PAIR, 0, t); add_inc(pid[t].y, T_PAIR, 1, t);
        }
        // PME grid, twiddles and LIMA's influence function (PME.cuh:434-495) in double
        m.elec = (elec && ncharged > 0) ? 1 : 0;
        for (int d = 0; d < 3; d++) {
            m.K[d] = p->pme_grid[d] > 0 ? p->pme_grid[d] : derive_grid_b(L[d], spacing);
            if (m.K[d] < 4 || m.K[d] > KMAX) throw LimaFail(LIMA_ERR_BAD_INPUT, "batched systems need PME grid edges in [4, 32]");
            m.tw0[d] = (int)twid.size();
            for (int k = 0; k < m.K[d]; k++) {
                const double ang = -2.0 * M_PI * (double)k / (double)m.K[d];
                twid.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
            }
            m.inv2h[d] = (float)(m.K[d] / (2.0 * L[d]));
        }
        m.H = m.K[0] / 2 + 1;
        m.NG = m.K[0] * m.K[1] * m.K[2];
        m.spec0 = spec_total;
        const int nspec = m.H * m.K[1] * m.K[2];
        spec_total += nspec;
        const double vcell = (L[0] / m.K[0]) * (L[1] / m.K[1]) * (L[2] / m.K[2]);
        m.fix_to_rho = (float)(1.0 / (FIX_SCALE * vcell));
        m.inv_ng = (float)(1.0 / (double)m.NG);
        {
            const double PI_L = (double)LIMA_PI_F, kappa = (double)B.pc.kappa, kc = (double)B.pc.kc;
            green.resize(spec_total);
            for (int z = 0; z < m.K[2]; z++)
                for (int y = 0; y < m.K[1]; y++)
                    for (int x = 0; x < m.H; x++) {
                        const int mi[3] = {x, y, z};
                        double k2 = 0, S = 1;
                        for (int d = 0; d < 3; d++) {
                            const int half = m.K[d] / 2;
                            const int mm = mi[d] <= half ? mi[d] : mi[d] - m.K[d];
                            const double kd = 2.0 * PI_L * mm / L[d];
                            k2 += kd * kd;
                            const double kh = kd * (L[d] / m.K[d]) * 0.5;
                            S *= (std::fabs(kh) < 1e-14) ? 1.0 : std::pow(std::sin(kh) / kh, 4);
                        }
                        double G = 0;
                        if (k2 > 1e-14) G = 4.0 * PI_L / k2 * std::exp(-k2 / (4.0 * kappa * kappa)) * S * S * kc;
                        green[m.spec0 + ((size_t)z * m.K[1] + y) * m.H + x] = (float)G;
                    }
            m.eself = -(double)B.pc.kappa / std::sqrt(PI_L) * qsq * kc;
        }
        if (!m.elec) m.eself = 0.0;
        m.lds_bytes = m.elec ? m.NG * 4 : 0;
        {
            const int64_t nt = (m.n + 63) / 64;
            if (part_total + nt * (nt + 1) / 2 * 128 >= (1ll << 31)) throw LimaFail(LIMA_ERR_BAD_INPUT, "batch too large");
            m.part0 = (int)part_total;
            part_total += nt * (nt + 1) / 2 * 128;
        }
    }
    B.N = (int64_t)u.size();
    // incidence CSR
    std::vector<int> inc_off(B.N + 1, 0), inc;
    for (int64_t i = 0; i < B.N; i++) {
        inc_off[i + 1] = inc_off[i] + (int)incl[i].size();
        inc.insert(inc.end(), incl[i].begin(), incl[i].end());
    }
    if (inc.empty()) inc.push_back(0);
    B.u_init = u;
    B.d_meta = dcopy(B.meta);
    B.u0 = dcopy(u);
    B.u_start = dcopy(u);
    HIP_CHECK(hipMalloc(&B.u1, B.N * sizeof(int4)));
    HIP_CHECK(hipMalloc(&B.f0, B.N * sizeof(float4)));
    HIP_CHECK(hipMalloc(&B.f1, B.N * sizeof(float4)));
    HIP_CHECK(hipMemset(B.f0, 0, B.N * sizeof(float4)));
    B.prm4 = dcopy(prm);
    B.comp = dcopy(comp);
    B.cinfo = dcopy(cinfo);
    B.cbm = dcopy(cbm);
    B.clut = dcopy(clut);
    B.inc_off = dcopy(inc_off);
    B.inc = dcopy(inc);
    B.bid = dcopy(bid); B.bp = dcopy(bp);
    B.aid = dcopy(aid); B.ap = dcopy(ap);
    B.did = dcopy(did); B.dp = dcopy(dp);
    B.iid = dcopy(iid); B.ip = dcopy(ip);
    B.pid = dcopy(pid); B.pp = dcopy(pp);
    B.green = dcopy(green);
    HIP_CHECK(hipMalloc(&B.spec, std::max(spec_total, 1) * sizeof(float2)));
    HIP_CHECK(hipMalloc(&B.spec2, std::max(spec_total, 1) * sizeof(float2)));
    HIP_CHECK(hipMalloc(&B.part, std::max<int64_t>(part_total, 1) * sizeof(float4)));
    B.twid = dcopy(twid);
    HIP_CHECK(hipMalloc(&B.out, nsys * sizeof(SysOut)));
    HIP_CHECK(hipMemset(B.out, 0, nsys * sizeof(SysOut)));
    HIP_CHECK(hipMalloc(&B.d_order, nsys * sizeof(int)));
    for (auto& st : B.streams) HIP_CHECK(hipStreamCreate(&st));
    // default shapes until autotuned: 1024 threads, two atoms per thread per pass
    B.tuning.n_buckets = NBUCKET;
    for (int k = 0; k < NBUCKET; k++) {
        B.tuning.bucket_max_atoms[k] = BUCKET_MAX[k];
        B.tuning.threads[k] = 1024;
        B.tuning.pair_unroll[k] = 2;
        B.tuning.us_per_iter[k] = 0.0;
    }
}

int shape_index(int threads, int rb) {
    for (int i = 0; i < NSHAPES; i++) if (SHAPES[i].threads == threads && SHAPES[i].unroll == rb) return i;
    return -1;
}

int bucket_of(const lima_batch_tuning& t, int n) {
    for (int k = 0; k < t.n_buckets; k++) if (n <= t.bucket_max_atoms[k]) return k;
    return t.n_buckets - 1;
}

BatchArgs make_args(const Batch& B, double tol_j, int max_iter, double h0) {
    BatchArgs a{};
    a.meta = B.d_meta; a.order = B.d_order;
    a.u0 = B.u0; a.u1 = B.u1; a.f0 = B.f0; a.f1 = B.f1;
    a.prm = B.prm4; a.comp = B.comp; a.cinfo = B.cinfo; a.cbm = B.cbm; a.clut = B.clut;
    a.inc_off = B.inc_off; a.inc = B.inc;
    a.bid = B.bid; a.bp = B.bp; a.aid = B.aid; a.ap = B.ap; a.did = B.did; a.dp = B.dp; a.iid = B.iid; a.ip = B.ip;
    a.pid = B.pid; a.pp = B.pp;
    a.green = B.green; a.spec = B.spec; a.spec2 = B.spec2; a.part = B.part; a.twid = B.twid; a.out = B.out;
    a.ctab = B.ctab; a.ctab_scale = (float)CTAB_N / B.rc;
    a.pc = B.pc; a.rc = B.rc; a.rc2 = B.rc * B.rc;
    a.tol = tol_j; a.h0 = h0; a.max_iter = max_iter;
    a.mode = MODE_PLAIN;
    a.feat = B.feat; a.scale = B.scale; a.feat_interval = 1; a.wlds = 0; a.hnext = B.hnext;
    a.rec = B.rec; a.rec_info = B.rec_info; a.rec_stride = B.rec_stride; a.rec_cap = B.rec_cap;
    return a;
}

void ensure_features(Batch& B) {
    if (!B.hnext) HIP_CHECK(hipMalloc(&B.hnext, B.N * 7 * sizeof(unsigned)));
    if (!B.feat) {
        HIP_CHECK(hipMalloc(&B.feat, B.N * 32 * sizeof(unsigned short)));
        HIP_CHECK(hipMemset(B.feat, 0, B.N * 32 * sizeof(unsigned short)));
    }
    if (!B.scale) HIP_CHECK(hipMalloc(&B.scale, B.N * sizeof(float)));
    for (int s = 0; s < B.nsys; s++)
        if (B.meta[s].n > HIST_MAX_ATOMS)
            throw LimaFail(LIMA_ERR_BAD_INPUT, "learned proposals and recording support systems of up to 4096 atoms");
}

// Launch the given systems, grouped by shape (one launch per shape, each on its own stream,
// largest systems first), and wait for all of them.
void launch_groups(Batch& B, const std::vector<int>& systems, const std::vector<int>& shape_of, const BatchArgs& a) {
    std::vector<int> order;
    std::vector<std::pair<int, int>> ranges(NSHAPES, {0, 0});
    for (int sh = 0; sh < NSHAPES; sh++) {
        std::vector<int> g;
        for (int s : systems) if (shape_of[s] == sh) g.push_back(s);
        std::stable_sort(g.begin(), g.end(), [&](int x, int y) { return B.meta[x].n > B.meta[y].n; });
        ranges[sh] = {(int)order.size(), (int)g.size()};
        order.insert(order.end(), g.begin(), g.end());
    }
    HIP_CHECK(hipMemcpy(B.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    for (int sh = 0; sh < NSHAPES; sh++) {
        const int cnt = ranges[sh].second;
        if (cnt == 0) continue;
        int lds = SHAPES[sh].threads / 64 * 192 * (int)sizeof(float);
        for (int k = 0; k < cnt; k++) {
            const SysMeta& m = B.meta[order[ranges[sh].first + k]];
            lds = std::max(lds, m.lds_bytes);
            if (a.mode != MODE_PLAIN) lds = std::max(lds, m.n * 28);   // histogram byte counters
            if (a.mode == MODE_LEARNED)   // counters after the j-force slices, in the pair pass
                lds = std::max(lds, SHAPES[sh].threads / 64 * 192 * (int)sizeof(float) + m.n * 28);
        }
        BatchArgs ag = a;
        if (a.mode == MODE_LEARNED) {   // the model's BF16 weights after everything else, for the whole launch
            ag.wlds = (lds + 15) & ~15;
            lds = ag.wlds + (MLP_H * MLP_IN + MLP_H * MLP_H) * (int)sizeof(unsigned short);
        }
        ag.order = B.d_order + ranges[sh].first;
        KernelFn fn = shape_kernel(sh);
        HIP_CHECK(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        hipLaunchKernelGGL(fn, dim3(cnt), dim3(SHAPES[sh].threads), (size_t)lds, B.streams[sh], ag);
        HIP_CHECK(hipGetLastError());
    }
    for (int sh = 0; sh < NSHAPES; sh++) HIP_CHECK(hipStreamSynchronize(B.streams[sh]));
}

std::vector<int> shapes_from_tuning(const Batch& B) {
    std::vector<int> shape_of(B.nsys);
    for (int s = 0; s < B.nsys; s++) {
        const int k = bucket_of(B.tuning, B.meta[s].n);
        int sh = shape_index(B.tuning.threads[k], B.tuning.pair_unroll[k]);
        if (sh < 0) sh = shape_index(1024, 2);
        shape_of[s] = sh;
    }
    return shape_of;
}

}  // namespace
}  // namespace lima

using lima::Batch;
using lima::LimaFail;
using lima::HipError;

struct lima_batch {
    Batch b;
};

namespace {
template <class F>
lima_status bguard(F&& f) {
    try {
        f();
        return LIMA_OK;
    } catch (const LimaFail& x) {
        lima::g_err = x.what();
        return x.code;
    } catch (const HipError& x) {
        lima::g_err = x.what();
        return LIMA_ERR_HIP;
    } catch (const std::exception& x) {
        lima::g_err = x.what();
        return LIMA_ERR_BAD_INPUT;
    }
}
}  // namespace

extern "C" {

lima_status lima_batch_create(const lima_box_desc* boxes, int32_t n_systems, const lima_simparams* params, lima_batch** out) {
    if (!out) { lima::g_err = "null output handle"; return LIMA_ERR_BAD_INPUT; }
    *out = nullptr;
    lima_batch* h = new lima_batch();
    const lima_status st = bguard([&] { lima::batch_create(h->b, boxes, n_systems, params); });
    if (st != LIMA_OK) {
        lima::batch_free(h->b);
        delete h;
        return st;
    }
    *out = h;
    return LIMA_OK;
}

lima_status lima_batch_destroy(lima_batch* h) {
    if (!h) return LIMA_OK;
    lima::batch_free(h->b);
    delete h;
    return LIMA_OK;
}

lima_status lima_batch_autotune(lima_batch* h, int32_t trial_iterations, lima_batch_tuning* out) {
    if (!h) { lima::g_err = "null batch"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        Batch& B = h->b;
        using namespace lima;
        const int iters = std::max(1, (int)trial_iterations);
        // per bucket: up to 8 systems spread over the bucket's size range
        std::vector<std::vector<int>> members(NBUCKET);
        for (int s = 0; s < B.nsys; s++) members[bucket_of(B.tuning, B.meta[s].n)].push_back(s);
        std::vector<int> sample;
        for (auto& g : members) {
            std::sort(g.begin(), g.end(), [&](int x, int y) { return B.meta[x].n < B.meta[y].n; });
            const int take = std::min<int>(8, (int)g.size());
            for (int k = 0; k < take; k++) sample.push_back(g[(size_t)k * g.size() / take]);
        }
        std::vector<std::vector<double>> cost(NSHAPES, std::vector<double>(NBUCKET, 0.0));
        std::vector<SysOut> res(B.nsys);
        for (int sh = 0; sh < NSHAPES; sh++) {
            std::vector<int> shape_of(B.nsys, sh);
            HIP_CHECK(hipMemcpy(B.u0, B.u_start, B.N * sizeof(int4), hipMemcpyDeviceToDevice));
            // tolerance 0: every system runs exactly `iters` trial steps
            launch_groups(B, sample, shape_of, make_args(B, 0.0, iters, 0.01));
            HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
            for (int s : sample) {
                const double us = (double)(res[s].t1 - res[s].t0) * 0.01 / (double)(res[s].iterations + 1);
                cost[sh][bucket_of(B.tuning, B.meta[s].n)] += us;
            }
        }
        for (int k = 0; k < NBUCKET; k++) {
            if (members[k].empty()) continue;
            int best = shape_index(1024, 2);
            for (int sh = 0; sh < NSHAPES; sh++) if (cost[sh][k] < cost[best][k]) best = sh;
            const int take = std::min<int>(8, (int)members[k].size());
            B.tuning.threads[k] = SHAPES[best].threads;
            B.tuning.pair_unroll[k] = SHAPES[best].unroll;
            B.tuning.us_per_iter[k] = cost[best][k] / take;
        }
        HIP_CHECK(hipMemcpy(B.u0, B.u_start, B.N * sizeof(int4), hipMemcpyDeviceToDevice));
        B.tuned = true;
        B.forces_valid = false;
        if (out) *out = B.tuning;
    });
}

lima_status lima_batch_set_tuning(lima_batch* h, const lima_batch_tuning* t) {
    if (!h || !t) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    if (t->n_buckets < 1 || t->n_buckets > 8) { lima::g_err = "n_buckets out of [1, 8]"; return LIMA_ERR_BAD_INPUT; }
    for (int k = 0; k < t->n_buckets; k++)
        if (lima::shape_index(t->threads[k], t->pair_unroll[k]) < 0) {
            lima::g_err = "unsupported shape (threads 256/512/1024, pair_unroll 1/2/4)";
            return LIMA_ERR_BAD_INPUT;
        }
    h->b.tuning = *t;
    h->b.tuned = true;
    return LIMA_OK;
}

lima_status lima_batch_minimize(lima_batch* h, double fmax_tol_kj, int64_t max_iter, double initial_step_nm, lima_em_result* results,
                                double* wall_seconds) {
    if (!h) { lima::g_err = "null batch"; return LIMA_ERR_BAD_INPUT; }
    if (max_iter < 0 || max_iter > (1ll << 30)) { lima::g_err = "max_iter out of range"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        std::vector<int> all(B.nsys);
        for (int s = 0; s < B.nsys; s++) all[s] = s;
        const std::vector<int> shape_of = shapes_from_tuning(B);
        const auto t0 = std::chrono::steady_clock::now();
        launch_groups(B, all, shape_of, make_args(B, fmax_tol_kj * 1000.0, (int)max_iter, initial_step_nm));
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        B.forces_valid = true;
        bool bad = false;
        for (int s = 0; s < B.nsys; s++) {
            if (!std::isfinite(res[s].e_acc)) bad = true;
            if (results) {
                results[s].iterations = res[s].iterations;
                results[s].accepted = res[s].accepted;
                results[s].final_energy = res[s].e_acc;
                results[s].final_fmax = (double)res[s].fmax / 1000.0;
                results[s].converged = res[s].converged;
                results[s].wall_seconds = (double)(res[s].t1 - res[s].t0) * 1e-8;
            }
        }
        if (wall_seconds) *wall_seconds = wall;
        if (bad) throw LimaFail(LIMA_ERR_NONFINITE, "non-finite energy in batched minimisation");
    });
}

lima_status lima_batch_minimize_learned(lima_batch* h, lima_mlp* model, double fmax_tol_kj, int64_t max_iter, double initial_step_nm,
                                        int32_t feature_interval, int32_t grow_on_learned, lima_em_result* results,
                                        double* wall_seconds) {
    if (!h || !model) { lima::g_err = "null batch or model"; return LIMA_ERR_BAD_INPUT; }
    if (max_iter < 0 || max_iter > (1ll << 30) || feature_interval < 1) {
        lima::g_err = "max_iter out of range or feature_interval < 1";
        return LIMA_ERR_BAD_INPUT;
    }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        ensure_features(B);
        std::vector<int> all(B.nsys);
        for (int s = 0; s < B.nsys; s++) all[s] = s;
        const std::vector<int> shape_of = shapes_from_tuning(B);
        BatchArgs ag = make_args(B, fmax_tol_kj * 1000.0, (int)max_iter, initial_step_nm);
        ag.mode = MODE_LEARNED;
        ag.mlp = mlp_view(model);
        ag.feat_interval = feature_interval;
        ag.grow_on_learned = grow_on_learned ? 1 : 0;
        const auto t0 = std::chrono::steady_clock::now();
        launch_groups(B, all, shape_of, ag);
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        B.forces_valid = true;
        bool bad = false;
        for (int s = 0; s < B.nsys; s++) {
            if (!std::isfinite(res[s].e_acc)) bad = true;
            if (results) {
                results[s].iterations = res[s].iterations;
                results[s].accepted = res[s].accepted;
                results[s].final_energy = res[s].e_acc;
                results[s].final_fmax = (double)res[s].fmax / 1000.0;
                results[s].converged = res[s].converged;
                results[s].wall_seconds = (double)(res[s].t1 - res[s].t0) * 1e-8;
            }
        }
        if (wall_seconds) *wall_seconds = wall;
        if (bad) throw LimaFail(LIMA_ERR_NONFINITE, "non-finite energy in batched minimisation");
    });
}

lima_status lima_batch_record(lima_batch* h, double fmax_tol_kj, int64_t max_iter, double initial_step_nm, int32_t atom_stride,
                              int32_t max_states, int64_t* n_states) {
    if (!h) { lima::g_err = "null batch"; return LIMA_ERR_BAD_INPUT; }
    if (max_iter < 0 || max_iter > (1ll << 30) || atom_stride < 1 || max_states < 1) {
        lima::g_err = "bad recording parameters";
        return LIMA_ERR_BAD_INPUT;
    }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        ensure_features(B);
        // per system: sampled atoms 0, stride, 2 stride, ...; max_states recorded accepted states
        B.h_rec_info.resize(B.nsys);
        int64_t off = 0;
        for (int s = 0; s < B.nsys; s++) {
            const int ns = (B.meta[s].n + atom_stride - 1) / atom_stride;
            B.h_rec_info[s] = make_int2((int)off, ns);
            off += (int64_t)ns * max_states * REC_W;
            if (off >= (1ll << 31)) throw LimaFail(LIMA_ERR_BAD_INPUT, "recording buffer exceeds 2^31 floats; raise the stride");
        }
        if (B.rec) { (void)hipFree(B.rec); B.rec = nullptr; }
        if (B.rec_info) { (void)hipFree(B.rec_info); B.rec_info = nullptr; }
        HIP_CHECK(hipMalloc(&B.rec, std::max<int64_t>(off, 1) * sizeof(float)));
        B.rec_info = dcopy(B.h_rec_info);
        B.rec_floats = off;
        B.rec_stride = atom_stride;
        B.rec_cap = max_states;
        std::vector<int> all(B.nsys);
        for (int s = 0; s < B.nsys; s++) all[s] = s;
        const std::vector<int> shape_of = shapes_from_tuning(B);
        BatchArgs ag = make_args(B, fmax_tol_kj * 1000.0, (int)max_iter, initial_step_nm);
        ag.mode = MODE_RECORD;
        launch_groups(B, all, shape_of, ag);
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        B.forces_valid = true;
        int64_t tot = 0;
        for (int s = 0; s < B.nsys; s++) tot += res[s].rec_iters;
        if (n_states) *n_states = tot;
    });
}

lima_status lima_batch_training_set(lima_batch* h, int32_t lookahead, float* features, float* targets, int64_t capacity,
                                    int64_t* n_samples) {
    if (!h || !n_samples || lookahead < 1) { lima::g_err = "bad argument"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        if (!B.rec) throw LimaFail(LIMA_ERR_BAD_INPUT, "nothing recorded: call lima_batch_record first");
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        // copy only the recorded states of each system
        std::vector<float> rec(B.rec_floats);
        for (int s = 0; s < B.nsys; s++) {
            const int2 ri = B.h_rec_info[s];
            const int64_t used = (int64_t)res[s].rec_iters * ri.y * REC_W;
            if (used > 0) HIP_CHECK(hipMemcpy(rec.data() + ri.x, B.rec + ri.x, used * sizeof(float), hipMemcpyDeviceToHost));
        }
        // target: the secant estimate of the atom's best step along its actual displacement D from
        // this accepted state to the one `lookahead` accepted states later.  With f, f' the force
        // components along D before and after, the force along D vanishes at |D| f / (f - f'); in
        // units of the atom's plain step |d| that is the ratio rho (soft atoms with small forces:
        // rho >> 1; stiff or overshooting ones: rho < 1).  rho in [0.1, 1000] maps log-linearly onto
        // the model's range [0.05, 3.95] (s ~ rho^0.47 up to a constant, which the proposal's
        // normalisation removes).
        int64_t cnt = 0;
        for (int s = 0; s < B.nsys; s++) {
            const int2 ri = B.h_rec_info[s];
            const int nst = res[s].rec_iters;
            for (int st = 0; st + lookahead < nst; st++)
                for (int q = 0; q < ri.y; q++) {
                    const float* e0 = &rec[ri.x + ((int64_t)st * ri.y + q) * REC_W];
                    const float* e1 = &rec[ri.x + ((int64_t)(st + lookahead) * ri.y + q) * REC_W];
                    const double d2 = (double)e0[32] * e0[32] + (double)e0[33] * e0[33] + (double)e0[34] * e0[34];
                    if (d2 < 1e-14) continue;   // atoms the step does not move carry no signal
                    double D[3], DD = 0.0;
                    for (int c = 0; c < 3; c++) {
                        int32_t u0, u1;
                        std::memcpy(&u0, &e0[35 + c], 4);
                        std::memcpy(&u1, &e1[35 + c], 4);
                        D[c] = (double)(u1 - u0) * 1e-8;
                        DD += D[c] * D[c];
                    }
                    if (DD < 1e-16) continue;
                    const double nD = std::sqrt(DD);
                    double f0 = 0.0, f1 = 0.0;
                    for (int c = 0; c < 3; c++) { f0 += e0[38 + c] * D[c] / nD; f1 += e1[38 + c] * D[c] / nD; }
                    double rho;
                    if (f0 <= 0.0) rho = 0.1;
                    else if (f0 - f1 <= 0.0) rho = 1000.0;
                    else rho = (f0 / (f0 - f1)) * nD / std::sqrt(d2);
                    const double u = std::min(1.0, std::max(0.0, (std::log10(std::max(rho, 1e-30)) + 1.0) / 4.0));
                    const double t = 0.05 * std::pow(3.95 / 0.05, u);
                    if (features && targets && cnt < capacity) {
                        std::memcpy(&features[cnt * 32], e0, 32 * sizeof(float));
                        targets[cnt] = (float)t;
                    }
                    cnt++;
                }
        }
        *n_samples = cnt;
    });
}

lima_status lima_batch_get_features(lima_batch* h, int32_t system, float* out) {
    if (!h || !out) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    if (system < 0 || system >= h->b.nsys) { lima::g_err = "system index out of range"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        if (!B.feat) throw LimaFail(LIMA_ERR_BAD_INPUT, "no features computed yet");
        const SysMeta& m = B.meta[system];
        std::vector<unsigned short> f((size_t)m.n * 32);
        HIP_CHECK(hipMemcpy(f.data(), B.feat + (int64_t)m.atom0 * 32, f.size() * sizeof(unsigned short), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < f.size(); k++) {
            const uint32_t bits = (uint32_t)f[k] << 16;
            std::memcpy(&out[k], &bits, 4);
        }
    });
}

lima_status lima_batch_get_counters(lima_batch* h, int64_t* out) {
    if (!h || !out) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        for (int s = 0; s < B.nsys; s++) {
            out[s * 3 + 0] = res[s].evaluations;
            out[s * 3 + 1] = res[s].learned_accepted;
            out[s * 3 + 2] = res[s].rec_iters;
        }
    });
}

lima_status lima_batch_get_state(lima_batch* h, int32_t system, lima_state_out* out) {
    if (!h || !out) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    if (system < 0 || system >= h->b.nsys) { lima::g_err = "system index out of range"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        const SysMeta& m = B.meta[system];
        std::vector<int4> u(m.n);
        std::vector<float4> f(m.n);
        HIP_CHECK(hipMemcpy(u.data(), B.u0 + m.atom0, m.n * sizeof(int4), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(f.data(), B.f0 + m.atom0, m.n * sizeof(float4), hipMemcpyDeviceToHost));
        if (!B.forces_valid) std::fill(f.begin(), f.end(), make_float4(0.f, 0.f, 0.f, 0.f));
        const int nC = B.h_ncomp[system];
        auto comp_u = [&](int i, int d) -> int64_t { return d == 0 ? u[i].x : (d == 1 ? u[i].y : u[i].z); };
        int first = 0;
        for (int c = 0; c < nC; c++) {
            const int nc = B.h_cn[system][c];
            const int k = first + B.h_key[system][c];
            int64_t base[3];
            for (int d = 0; d < 3; d++) {
                const int64_t wk = comp_u(k, d) - fdiv64(comp_u(k, d), m.Lq[d]) * m.Lq[d];
                const int32_t org = (int32_t)(wk / 100000000ll);
                if (out->compound_origo) out->compound_origo[c * 3 + d] = org;
                base[d] = comp_u(k, d) - wk + (int64_t)org * 100000000ll;
            }
            for (int sl = 0; sl < nc; sl++) {
                const int i = first + sl;
                const size_t ps = (size_t)c * 32 + sl;
                const float ff[3] = {f[i].x, f[i].y, f[i].z};
                for (int d = 0; d < 3; d++) {
                    if (out->particle_coord) out->particle_coord[ps * 3 + d] = (int32_t)(comp_u(i, d) - base[d]);
                    if (out->particle_velocity) out->particle_velocity[ps * 3 + d] = 0.f;
                    if (out->particle_force) out->particle_force[ps * 3 + d] = ff[d];
                }
                if (out->particle_pot_energy) out->particle_pot_energy[ps] = 0.f;
            }
            first += nc;
        }
        for (int t = 0; t < B.h_ntiny[system]; t++) {
            const int i = first + t;
            for (int d = 0; d < 3; d++) {
                const int64_t ww = comp_u(i, d) - fdiv64(comp_u(i, d), m.Lq[d]) * m.Lq[d];
                const int32_t blk = (int32_t)(ww / 100000000ll);
                if (out->tinymol_block) out->tinymol_block[t * 3 + d] = blk;
                if (out->tinymol_coord) out->tinymol_coord[t * 3 + d] = (int32_t)(ww - (int64_t)blk * 100000000ll);
                if (out->tinymol_velocity) out->tinymol_velocity[t * 3 + d] = 0.f;
                if (out->tinymol_force) out->tinymol_force[t * 3 + d] = (&f[i].x)[d];
            }
            if (out->tinymol_pot_energy) out->tinymol_pot_energy[t] = 0.f;
        }
    });
}

lima_status lima_batch_reset(lima_batch* h) {
    if (!h) { lima::g_err = "null batch"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        Batch& B = h->b;
        HIP_CHECK(hipMemcpy(B.u0, B.u_start, B.N * sizeof(int4), hipMemcpyDeviceToDevice));
        B.forces_valid = false;
    });
}

lima_status lima_batch_get_phase_times(lima_batch* h, double* seconds) {
    if (!h || !seconds) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return bguard([&] {
        using namespace lima;
        Batch& B = h->b;
        std::vector<SysOut> res(B.nsys);
        HIP_CHECK(hipMemcpy(res.data(), B.out, B.nsys * sizeof(SysOut), hipMemcpyDeviceToHost));
        for (int s = 0; s < B.nsys; s++)
            for (int q = 0; q < 7; q++) seconds[s * 7 + q] = (double)res[s].ph[q] * 1e-8;
    });
}

lima_status lima_batch_get_tuning(lima_batch* h, lima_batch_tuning* out) {
    if (!h || !out) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    *out = h->b.tuning;
    return LIMA_OK;
}

}  // extern "C"
