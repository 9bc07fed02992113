// forces.hip -- nonbonded and bonded force kernels (FP32 arithmetic, FP64 energy totals).
//
// Pair kernel: LIMA's cutoff pair path (compoundFarneighborShortrangeInteractionsKernel,
// code/LIMA_ENGINE/src/EngineKernels.cuh:47-146 with LennardJonesInteractions.cuh:164-203).
// A wave owns one i-cluster (8 particles) and walks its entry list 8 entries at a time:
// lane = 8*jj + ii computes pair (i-slot ii, entry jj), so one wave instruction covers 64 pairs.
// The list is full (both sides of every pair, as LIMA evaluates them): only i forces are
// accumulated, so the loop has no atomics; a cluster split over several waves of one workgroup
// sums their shares in LDS in a fixed order, and one wave stores the result (deterministic).
// FP32 VALU bound: about 50 VALU slots per evaluated pair (pair_scalar_cut in lima_dev.h).
//
// Intra kernel: LIMA's immediate-neighbour path (EngineKernels.cuh:150-263,
// LennardJonesInteractions.cuh:121-161): pairs of the same or bonded compounds, BondedParticlesLUT
// exclusions, no cutoff, force on self only (as LIMA evaluates it); one wave per compound walks its
// items (itself and its bonded compounds) and stores the compound's forces once.  Within a compound the
// displacement is unwrapped; a bonded compound is first moved to the image whose key particle is
// nearest (LIMA's hyperpos of the neighbour's origo, EngineKernels.cuh:235-236).
//
// Bonded kernel: ForceComputations.cuh:17-264,335-376; one lane per (atom, term) incidence, the
// lanes of an atom summed in-wave and added to its cluster slot once.
#include "engine.h"
#include "lima_terms.h"

namespace lima {

namespace {

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Slot k of the cluster list: position relative to the cluster reference point and charge, and
// (tiled PME) the charged particle put in its tile bin.  nslots: the slot capacity;
// slots of clusters beyond the current count are left alone.  The force buffers need no clearing:
// every kernel overwrites the slots it owns.
__global__ void k_pack(const int* __restrict__ ncl_dev, int nslots, const int* __restrict__ catoms, const uint4* __restrict__ w,
                       const float4* __restrict__ prm,
                       const uint4* __restrict__ cref, long long Lqx, long long Lqy, long long Lqz,
                       float4* __restrict__ xq, BinArgs bins) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = k < min(*ncl_dev * CLUSTER, nslots);
    const int a = live ? catoms[k] : -1;
    uint4 p = make_uint4(0u, 0u, 0u, 0u);
    float q = 0.f;
    if (live) {
        float4 out = make_float4(PAD_POS, PAD_POS, PAD_POS, 0.f);
        if (a >= 0) {
            p = w[a];
            const uint4 r = cref[k >> 3];
            long long d[3] = {(long long)p.x - (long long)r.x, (long long)p.y - (long long)r.y, (long long)p.z - (long long)r.z};
            const long long L[3] = {Lqx, Lqy, Lqz};
            for (int i = 0; i < 3; i++) {
                if (2 * d[i] > L[i]) d[i] -= L[i];
                else if (2 * d[i] < -L[i]) d[i] += L[i];
            }
            q = prm[a].x;
            out = make_float4((float)d[0] * LIMA_LIMA_TO_NANO_F, (float)d[1] * LIMA_LIMA_TO_NANO_F, (float)d[2] * LIMA_LIMA_TO_NANO_F, q);
        }
        xq[k] = out;
    }
    if (bins.bin_count) bin_wave(bins, a >= 0 && q != 0.f, k, p, q);
}

struct PairArgs {
    const int* ncl_dev;  // clusters of the current list (device)
    int ncl_max, cap;
    int split;           // waves per i-cluster (each takes a contiguous share of its entries)
    float rc2;
    float zscale;        // COUL_TAB / cutoff^2
    PairConst pc;
    const float4* coul_tab;
    const float4* xq;
    const float2* ljs;
    const int* ecount;   // near entries of each cluster (front of its slice)
    const int* efar;     // far entries (back of its slice)
    const uint2* ent;    // list entries: x = j slot | mask << 24, y = shift code | -1 / +1 box bits of axis d at 8 + 2d / 9 + 2d
    const uint4* cref;   // cluster reference points (integer)
    uint32_t lq[3];      // box lengths, integer units (mod 2^32)
    float4* fpair;   // [nslots] (fx, fy, fz, potE share)
    double* epart;   // this kernel's energy region: one record per wave
};

// One i-cluster per `split` waves: lane = 8 jj + ii computes pair (i slot ii, entry jj) of 8
// consecutive entries.  The list is full (every pair appears from both sides), so only the i
// force is accumulated: no atomics and no cross-lane traffic inside the loop, and the pair
// forces are deterministic.
//
// Entries are staged through LDS 64 at a time: lane l loads entry l of the chunk, its j particle
// and the j cluster's reference point, forms the shift between the two clusters' reference
// points (an exact integer difference, rounded once to float: the value the list build would
// compute), adds it to the j position (stored relative to its own reference point) and stores
// (x, y, z, mask bits) and (sigma/2, sqrt(24 eps), q) in the wave's LDS slice.  The 8 lanes of a
// group then read their entry with broadcast LDS reads, the first record in every pass and the
// second only in passes with a lane inside the cutoff, so the texture path sees each datum once
// instead of eight times.  Two LDS buffers: chunk c + 1 is loaded while c is computed.
// A cluster's near entries come first, its far entries (beyond the cutoff from its box at the
// build) last.  Before an entry is staged, the gap between its j position and the bounding box of
// the cluster's own particles at this step is compared with the cutoff: an entry farther away
// cannot reach any i, so it is dropped and the chunk's survivors are compacted (ballot prefix)
// into fewer passes.  The test is exact at every step, so the list's skin costs fetches, not
// passes.
constexpr int PAIR_WAVES = 4;
constexpr int CHUNK = 64;

template <bool ENERGY>
__global__ __launch_bounds__(256) void k_pair(PairArgs a) {
    __shared__ float4 sA[PAIR_WAVES][2][CHUNK];
    __shared__ float4 sB[PAIR_WAVES][2][CHUNK];
    __shared__ float4 tab[COUL_TAB + 1];
    __shared__ float4 part_f[PAIR_WAVES][CLUSTER];   // each wave's share of its cluster's i forces (+ potE)
    for (int i = threadIdx.x; i <= COUL_TAB; i += blockDim.x) tab[i] = a.coul_tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // a workgroup holds PAIR_WAVES / split whole clusters: the waves of one cluster are consecutive
    const int wv = xcd_block(blockIdx.x, gridDim.x) * PAIR_WAVES + wid;
    const int I = wv / a.split, part = wv - I * a.split;
    const bool live = I < min(*a.ncl_dev, a.ncl_max);   // the grid covers the capacity (uniform per wave)
    const int ii = lane & 7, jj = lane >> 3;
    const uint32_t ibit = 1u << ii;
    const int islot = I * CLUSTER + ii;
    float4 xi = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 pi = make_float2(0.f, 0.f);
    if (live) { xi = a.xq[islot]; pi = a.ljs[islot]; }
    const float qi_kc = xi.w * a.pc.kc;
    // bounding box of the cluster's real particles at this step (padding slots sit at PAD_POS):
    // centre and half extent per axis, and the cutoff with 1e-4 nm to spare for rounding
    float bc[3], bh[3];
    {
        const bool real = fabsf(xi.x) < 0.5f * PAD_POS;
        const float p3[3] = {xi.x, xi.y, xi.z};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float lo = -wave_max(real ? -p3[d] : -1e30f), hi = wave_max(real ? p3[d] : -1e30f);
            bc[d] = 0.5f * (lo + hi);
            bh[d] = 0.5f * (hi - lo);
        }
    }
    const float rcut = sqrtf(a.rc2) + 1e-4f, rcut2 = rcut * rcut;
    // this cluster's reference point; the shift of an entry is ref_J - ref_I + (p - n) L per axis
    // from its +1 / -1 box bits, in 32-bit wrap-around arithmetic, which is exact since every
    // wanted difference is far below 2^31 units.  Masks, not selects: a select between values
    // the lambdas capture by reference becomes a load through a selected pointer, which forces
    // them into scratch memory.
    uint4 ri = live ? a.cref[I] : make_uint4(0u, 0u, 0u, 0u);
    ri = make_uint4(__builtin_amdgcn_readfirstlane(ri.x), __builtin_amdgcn_readfirstlane(ri.y), __builtin_amdgcn_readfirstlane(ri.z), 0u);   // wave-uniform
    const uint32_t lqx = a.lq[0], lqy = a.lq[1], lqz = a.lq[2];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    double elj = 0.0, ecl = 0.0;   // FP64 energy accumulation (energy steps only)
    double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // -1/2 S d (x) d per evaluated pair (energy steps only)
    int npairs = 0;
    const int nn = live ? a.ecount[I] : 0, nf = live ? a.efar[I] : 0;
    const int ntot = (nn + nf) / CLUSTER;
    const int it0 = (int)((int64_t)ntot * part / a.split), it1 = (int)((int64_t)ntot * (part + 1) / a.split);
    const int64_t base = (int64_t)I * a.cap;
    const int v0 = it0 * CLUSTER;                    // this wave's first entry in the near + far order
    const int nent = (it1 - it0) * CLUSTER;          // entries of this wave (multiple of 8)
    const int nchunk = (nent + CHUNK - 1) / CHUNK;   // wave-uniform

    // raw loads of the chunk in flight: j position, LJ pair, j cluster reference point (xyz) and
    // the entry's shift code | mask << 24; the shift is formed in stage(), after the compute
    // that the loads overlap
    float4 gxj = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 gpj = make_float2(0.f, 0.f);
    uint4 grj = make_uint4(0u, 0u, 0u, 0u);
    auto fetch = [&](int c) {   // global loads of chunk c into registers (lane l: entry l)
        const int k = c * CHUNK + lane;
        if (k < nent) {
            const int v = v0 + k;
            const uint2 rec = a.ent[base + (v < nn ? v : a.cap - nf + (v - nn))];
            const int b = (int)(rec.x & 0xffffffu);
            const uint4 rj = a.cref[b / CLUSTER];
            gxj = a.xq[b];
            gpj = a.ljs[b];
            grj = make_uint4(rj.x, rj.y, rj.z, (rec.y & 0x3fffu) | (rec.x & 0xff000000u));
        } else {   // past the end: far away, mask 0, no shift
            gxj = make_float4(PAD_POS, PAD_POS, PAD_POS, 0.f);
            gpj = make_float2(0.f, 0.f);
            grj = make_uint4(ri.x, ri.y, ri.z, 13u);
        }
    };    // the fetched chunk's entries inside the cluster's reach, compacted into LDS buffer buf and padded
    // to whole passes; returns how many passes they fill (wave-uniform)
    auto stage = [&](int buf) {
        const uint32_t cw = grj.w;
        auto bitmask = [cw](int b) { return (uint32_t)((int32_t)(cw << (31 - b)) >> 31); };   // bit b: 0 or all ones
        const float shx = (float)(int)(grj.x - ri.x + (lqx & bitmask(9)) - (lqx & bitmask(8))) * LIMA_LIMA_TO_NANO_F;
        const float shy = (float)(int)(grj.y - ri.y + (lqy & bitmask(11)) - (lqy & bitmask(10))) * LIMA_LIMA_TO_NANO_F;
        const float shz = (float)(int)(grj.z - ri.z + (lqz & bitmask(13)) - (lqz & bitmask(12))) * LIMA_LIMA_TO_NANO_F;
        // (x, y, z, mask bits) for the test every pass makes, (sigma/2, sqrt(24 eps), q) for the force
        const float4 ga = make_float4(gxj.x + shx, gxj.y + shy, gxj.z + shz, __uint_as_float(cw >> 24));
        const float4 gb = make_float4(gpj.x, gpj.y, gxj.w, 0.f);
        const float gx = fmaxf(fabsf(ga.x - bc[0]) - bh[0], 0.f), gy = fmaxf(fabsf(ga.y - bc[1]) - bh[1], 0.f),
                    gz = fmaxf(fabsf(ga.z - bc[2]) - bh[2], 0.f);
        const bool keep = gx * gx + gy * gy + gz * gz <= rcut2;   // entries past the end sit at PAD_POS
        const uint64_t bal = __ballot(keep);
        const int cnt = __popcll(bal);
        const int pos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        if (keep) {
            sA[wid][buf][pos] = ga;
            sB[wid][buf][pos] = gb;
        }
        const int npad = ((cnt + CLUSTER - 1) & ~(CLUSTER - 1)) - cnt;
        if (lane < npad) {
            sA[wid][buf][cnt + lane] = make_float4(PAD_POS, PAD_POS, PAD_POS, __uint_as_float(0u));
            sB[wid][buf][cnt + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return (cnt + CLUSTER - 1) / CLUSTER;
    };
    int nit_next = 0;
    if (nchunk > 0) {
        fetch(0);
        nit_next = stage(0);
        __builtin_amdgcn_wave_barrier();   // the wave's own LDS slice: ordering only, no s_barrier
    }
    for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        const int nit = __builtin_amdgcn_readfirstlane(nit_next);   // wave-uniform
        if (c + 1 < nchunk) fetch(c + 1);   // in flight during the compute below
        for (int k = 0; k < nit; k++) {
            const float4 A = sA[wid][buf][k * CLUSTER + jj];
            const uint32_t m = __float_as_uint(A.w);
            const float dx = A.x - xi.x, dy = A.y - xi.y, dz = A.z - xi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const bool inc = (m & ibit) && (r2 <= a.rc2);   // the list masks exclude self and padding pairs
            // the force arithmetic runs under the lanes' own mask (no selects); a pass where no lane
            // is inside the cutoff branches over it
            if (inc) {
                if (ENERGY) npairs++;
                float el = 0.f, ec = 0.f;
                const float4 B = sB[wid][buf][k * CLUSTER + jj];
                const float S = pair_scalar_tab(r2, pi.x + B.x, pi.y * B.y, qi_kc * B.z, a.pc, tab, a.zscale,
                                                ENERGY ? &el : nullptr, &ec);
                fx = fmaf(dx, S, fx); fy = fmaf(dy, S, fy); fz = fmaf(dz, S, fz);
                if (ENERGY) {
                    elj += 0.5 * (double)el;   // each pair is evaluated from both sides: half per side
                    ecl += 0.5 * (double)ec;
                    const float hx = -0.5f * S * dx, hy = -0.5f * S * dy, hz = -0.5f * S * dz;
                    vir[0] += (double)(hx * dx); vir[1] += (double)(hy * dy); vir[2] += (double)(hz * dz);
                    vir[3] += (double)(hx * dy); vir[4] += (double)(hx * dz); vir[5] += (double)(hy * dz);
                }
            }
        }
        if (c + 1 < nchunk) nit_next = stage(buf ^ 1);
        __builtin_amdgcn_wave_barrier();
    }
    // i force: lanes ii, ii+8, ..., ii+56 hold partial sums
    for (int o = 8; o < 64; o <<= 1) {
        fx += __shfl_xor(fx, o, 64);
        fy += __shfl_xor(fy, o, 64);
        fz += __shfl_xor(fz, o, 64);
    }
    float ep = 0.f;
    if (ENERGY) {
        ep = (float)(elj + ecl);
        for (int o = 8; o < 64; o <<= 1) ep += __shfl_xor(ep, o, 64);
    }
    if (a.split == 1) {
        if (live && jj == 0) {
            put_force<ENERGY>(a.fpair, islot, fx, fy, fz, ep);
        }
    } else {   // the cluster's waves meet in LDS; its first wave adds the shares in part order
        if (jj == 0) part_f[wid][ii] = make_float4(fx, fy, fz, ep);
        __syncthreads();
        if (live && part == 0 && jj == 0) {
            float4 t = part_f[wid][ii];
            for (int q = 1; q < a.split; q++) {
                const float4 u = part_f[wid + q][ii];
                t.x += u.x; t.y += u.y; t.z += u.z; t.w += u.w;
            }
            put_force<ENERGY>(a.fpair, islot, t.x, t.y, t.z, t.w);
        }
    }
    if (ENERGY) {
        const double tl = wave_sum_d(elj), tc = wave_sum_d(ecl), np = wave_sum_d((double)npairs);
#pragma unroll
        for (int k = 0; k < 6; k++) vir[k] = wave_sum_d(vir[k]);
        if (lane == 0 && wv < a.ncl_max * a.split)   // waves past the live count write zero records
            write_record(a.epart + (int64_t)wv * ACC_N, tl, tc, 0.0, 0.0, np * 0.5, vir);   // pairs: both sides counted
    }
}

struct IntraArgs {
    int n_groups;
    long long Lq[3];
    PairConst pc;
    const int2* groups;   // (first item, item count) per compound A
    const int4* items;    // (A, B, key particle of A, key particle of B)
    const uint32_t* lut;
    const uint4* w;
    const int4* img;
    const float4* prm;
    const int* slot_of;
    const int64_t* first;
    float4* fintra;  // [nslots] (fx, fy, fz, potE share)
    double* epart;
};

__device__ __forceinline__ void unwrapped3(const uint4* w, const int4* img, const long long* Lq, int a, long long* u) {
    const uint4 p = w[a];
    const int4 m = img[a];
    u[0] = (long long)p.x + (long long)m.x * Lq[0];
    u[1] = (long long)p.y + (long long)m.y * Lq[1];
    u[2] = (long long)p.z + (long long)m.z * Lq[2];
}

// One wave per compound A: for each of A's items (A itself and its bonded compounds B), B's
// particles (unwrapped Coord, parameters) are staged in the wave's LDS slice, then lanes 0..31 take
// i = lane with j in [0,16) and lanes 32..63 the same i with j in [16,32).  A's forces accumulate
// in registers over its items and are stored once.
struct IntraLds {
    long long u[4][32][3];
    float4 p[4][32];
};

template <bool ENERGY>
__device__ __forceinline__ void intra_items(const IntraArgs& a, int blk, IntraLds& L) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int g = blk * 4 + wid;
    if (g >= a.n_groups) return;
    const int2 grp = a.groups[g];
    const int A = a.items[grp.x].x;
    const int nA = (int)(a.first[A + 1] - a.first[A]);
    const int i = lane & 31, half = lane >> 5;
    const int ai = (int)a.first[A] + i;
    long long ui[3] = {0, 0, 0};
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < nA) { unwrapped3(a.w, a.img, a.Lq, ai, ui); pi = a.prm[ai]; }
    float fx = 0.f, fy = 0.f, fz = 0.f, el_s = 0.f, ec_s = 0.f;
    float vir[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // -1/2 S d (x) d per evaluated pair (energy evaluations)
    for (int it = grp.x; it < grp.x + grp.y; it++) {
        const int4 item = a.items[it];
        const int B = item.y;
        const int nB = (int)(a.first[B + 1] - a.first[B]);
        long long sh[3] = {0, 0, 0};   // B's image: its key particle nearest to A's (LIMA's hyperpos)
        if (A != B) {
            long long ka[3], kb[3];
            unwrapped3(a.w, a.img, a.Lq, item.z, ka);
            unwrapped3(a.w, a.img, a.Lq, item.w, kb);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const long long dk = kb[d] - ka[d];
                long long r = dk % a.Lq[d];
                r += r < 0 ? a.Lq[d] : 0;
                r -= 2 * r > a.Lq[d] ? a.Lq[d] : 0;
                sh[d] = r - dk;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the previous item's reads are done
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < nB) {
            const int bj = (int)a.first[B] + lane;
            long long ub[3];
            unwrapped3(a.w, a.img, a.Lq, bj, ub);
            L.u[wid][lane][0] = ub[0] + sh[0];
            L.u[wid][lane][1] = ub[1] + sh[1];
            L.u[wid][lane][2] = ub[2] + sh[2];
            L.p[wid][lane] = a.prm[bj];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (i < nA) {
            const uint32_t row = a.lut[it * 32 + i];
            for (int j = half * 16; j < half * 16 + 16 && j < nB; j++) {
                if ((row >> j) & 1u) continue;
                const long long* uj = L.u[wid][j];
                const float dx = (float)(uj[0] - ui[0]) * LIMA_LIMA_TO_NANO_F;
                const float dy = (float)(uj[1] - ui[1]) * LIMA_LIMA_TO_NANO_F;
                const float dz = (float)(uj[2] - ui[2]) * LIMA_LIMA_TO_NANO_F;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (!(r2 > 0.f)) continue;
                const float4 pj = L.p[wid][j];
                float el = 0.f, ec = 0.f;
                // no cutoff here: pair_scalar_cut's erfc polynomial extrapolates beyond kappa r = 3.25 with an
                // absolute error below 1e-12 (erfc itself is < 2e-6 there)
                const float S = pair_scalar_cut(r2, pi.y + pj.y, 24.0f * (pi.z * pj.z), a.pc.kc * (pi.x * pj.x), a.pc,
                                                ENERGY ? &el : nullptr, &ec);
                fx += dx * S; fy += dy * S; fz += dz * S;
                if (ENERGY) {
                    el_s += 0.5f * el; ec_s += 0.5f * ec;
                    const float hx = -0.5f * S * dx, hy = -0.5f * S * dy, hz = -0.5f * S * dz;
                    vir[0] += hx * dx; vir[1] += hy * dy; vir[2] += hz * dz;
                    vir[3] += hx * dy; vir[4] += hx * dz; vir[5] += hy * dz;
                }
            }
        }
    }
    fx += __shfl_xor(fx, 32, 64);
    fy += __shfl_xor(fy, 32, 64);
    fz += __shfl_xor(fz, 32, 64);
    float e_i = 0.f;
    if (ENERGY) e_i = __shfl_xor(el_s + ec_s, 32, 64) + el_s + ec_s;
    if (half == 0 && i < nA) {
        const int s = a.slot_of[ai];
        put_force<ENERGY>(a.fintra, s, fx, fy, fz, e_i);
    }
    if (ENERGY) {
        const double tl = wave_sum_d((double)el_s), tc = wave_sum_d((double)ec_s);
        double vd[6];
#pragma unroll
        for (int k = 0; k < 6; k++) vd[k] = wave_sum_d((double)vir[k]);
        if (lane == 0) write_record(a.epart + (int64_t)g * ACC_N, tl, tc, 0.0, 0.0, 0.0, vd);
    }
}

struct BondArgs {
    int nb, na, nd, ni, np;
    long long Lq[3];
    PairConst pc;
    const int2* bid; const float2* bp;
    const int4* aid; const float4* ap;
    const int4* did; const float4* dp;
    const int4* iid; const float2* ip;
    const int2* pid; const float2* pp;
    int n_lanes;
    int multi_lane0;     // workgroups from this lane on hold one atom each, one run per wave
    const uint32_t* lane_code;
    const int* lane_atom;
    const uint4* w;
    const int4* img;
    const int* slot_of;
    float4* fb;      // [nslots] (fx, fy, fz, potE share)
    double* epart;
};

// x_to - x_from [nm] of two bonded particles.  Bonded displacements are far shorter than half the
// box, so the minimum image of the wrapped Coord difference is the same integer as the unwrapped
// difference LIMA uses (no image counters needed).
__device__ __forceinline__ Dq bdispq(const BondArgs& a, int from, int to) {
    const uint4 p0 = a.w[from], p1 = a.w[to];
    long long d[3] = {(long long)p1.x - (long long)p0.x, (long long)p1.y - (long long)p0.y, (long long)p1.z - (long long)p0.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (2 * d[k] > a.Lq[k]) d[k] -= a.Lq[k];
        else if (2 * d[k] < -a.Lq[k]) d[k] += a.Lq[k];
    }
    return {d[0], d[1], d[2]};
}
__device__ __forceinline__ v3 bdisp(const BondArgs& a, int from, int to) { return dq_nm(bdispq(a, from, to)); }

// Bonded terms, one lane per (atom, term) incidence.  An atom's incidences occupy consecutive
// lanes of one wave (at most 64, checked at creation); each lane evaluates its term
// (ForceComputations.cuh:17-376) and keeps the atom's share, a segmented shuffle sum gathers the
// run into its first lane, and that lane stores the atom's bonded force in the slot's bonded
// array (fb): one plain store per component, no atomics.  fb is cleared at every list rebuild, so
// slots of atoms without terms read zero.  Every lane holds one short dependency chain, so long
// incidence lists (chain atoms belong to ~10 terms) do not serialise; per-atom sums are in a
// fixed order.  LIMA splits a term's potE equally over its atoms (EngineKernels.cuh:776-790);
// the FP64 total counts each term once, at its role-0 lane.
template <bool ENERGY>
__device__ __forceinline__ void bonded_lanes(const BondArgs& a, int blk) {
    const int k = blk * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t code = LANE_PAD;
    int la = lane << 24;
    if (k < a.n_lanes) { code = a.lane_code[k]; la = a.lane_atom[k]; }
    const int last = (la >> 24) & 63, atom = la & 0xffffff;
    v3 f = {0.f, 0.f, 0.f};
    float e = 0.f, eshare = 0.f;
    float vir[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // (x_atom - x_ref) (x) f, symmetric part (energy evaluations)
    int role = -1;
    if (code != LANE_PAD) {
        const int type = (int)(code >> 29), t = (int)(code & (INC_MAX_TERMS - 1));
        role = (int)((code >> 27) & 3u);
        int nat, ref;   // ref: the term's role-0 particle, the origin of its virial positions
        if (type == T_BOND) {   // calcSinglebondForces :17-51
            const int2 id = a.bid[t];
            ref = id.x;
            v3 fa;
            e = term_bond(bdispq(a, id.y, id.x), a.bp[t], &fa);
            f = role == 0 ? fa : -fa;
            nat = 2;
        } else if (type == T_ANGLE) {   // calcAnglebondForces :53-115, harmonic in theta + Urey-Bradley
            const int4 id = a.aid[t];
            ref = id.x;
            v3 fl, fr;
            e = term_angle(bdisp(a, id.y, id.x), bdisp(a, id.y, id.z), bdispq(a, id.z, id.x), a.ap[t], &fl, &fr);
            f = role == 0 ? fl : (role == 2 ? fr : -(fl + fr));
            nat = 3;
        } else if (type == T_DIH) {   // calcDihedralbondForces :119-210, phi_L = -atan2(sin, cos)
            const int4 id = a.did[t];
            ref = id.x;
            v3 fd[4];
            e = term_dihedral(bdisp(a, id.x, id.y), bdisp(a, id.y, id.z), bdisp(a, id.z, id.w), a.dp[t], fd);
            f = pick_role4(fd, role);
            nat = 4;
        } else if (type == T_IMP) {   // calcImproperdihedralbondForces :214-264
            const int4 id = a.iid[t];
            ref = id.x;
            v3 fd[4];
            e = term_improper(bdisp(a, id.x, id.y), bdisp(a, id.x, id.z), bdisp(a, id.x, id.w), a.ip[t], fd);
            f = pick_role4(fd, role);
            nat = 4;
        } else {   // computePairbondForces :335-376
            const int2 id = a.pid[t];
            ref = id.x;
            v3 fa;
            e = term_pair14(bdisp(a, id.x, id.y), a.pp[t], &fa);
            f = role == 0 ? fa : -fa;
            nat = 2;
        }
        eshare = e / (float)nat;
        if (ENERGY && role != 0) {   // the role-0 lane sits at the origin
            const v3 r = bdisp(a, ref, atom);
            vir[0] = r.x * f.x; vir[1] = r.y * f.y; vir[2] = r.z * f.z;
            vir[3] = 0.5f * (r.x * f.y + r.y * f.x);
            vir[4] = 0.5f * (r.x * f.z + r.z * f.x);
            vir[5] = 0.5f * (r.y * f.z + r.z * f.y);
        }
    }
    // run sums into the run's first lane: after the step with offset o, lane i holds the sum of
    // lanes i .. min(i + 2o - 1, last)
    for (int o = 1; o < 64; o <<= 1) {
        const float ox = __shfl_down(f.x, o, 64), oy = __shfl_down(f.y, o, 64), oz = __shfl_down(f.z, o, 64);
        const float oe = ENERGY ? __shfl_down(eshare, o, 64) : 0.f;
        if (lane + o <= last) {
            f.x += ox; f.y += oy; f.z += oz;
            if (ENERGY) eshare += oe;
        }
    }
    if (blk * (int)blockDim.x < a.multi_lane0) {   // workgroup-uniform: every atom in one run
        const int prev = __shfl_up(la, 1, 64);
        if (code != LANE_PAD && (lane == 0 || prev != la)) {   // the run's first lane
            const int s = a.slot_of[atom];
            put_force<ENERGY>(a.fb, s, f.x, f.y, f.z, eshare);
        }
    } else {   // one atom, its runs starting at the waves' first lanes: the run sums meet in LDS
        __shared__ float4 runs[4];
        const int w = threadIdx.x >> 6;
        if (lane == 0) runs[w] = code != LANE_PAD ? make_float4(f.x, f.y, f.z, eshare) : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        if (threadIdx.x == 0 && code != LANE_PAD) {
            float4 t = runs[0];
            for (int r = 1; r < 4; r++) { t.x += runs[r].x; t.y += runs[r].y; t.z += runs[r].z; t.w += runs[r].w; }
            const int s = a.slot_of[atom];
            put_force<ENERGY>(a.fb, s, t.x, t.y, t.z, t.w);
        }
    }
    if (ENERGY) {
        const double tot = wave_sum_d(role == 0 ? (double)e : 0.0);
        double vd[6];
#pragma unroll
        for (int c = 0; c < 6; c++) vd[c] = wave_sum_d((double)vir[c]);
        if (lane == 0 && (k & ~63) < a.n_lanes)   // waves inside the region
            write_record(a.epart + (int64_t)(k >> 6) * ACC_N, 0.0, 0.0, tot, 0.0, 0.0, vd);
    }
}

// Intra-compound pairs and bonded terms in one launch: blocks [0, nib) take intra items (4 per
// block), the rest bonded terms (256 per block).  Both are short, latency-bound passes; sharing a
// launch lets their memory latencies overlap.
template <bool ENERGY>
__global__ __launch_bounds__(256) void k_local(IntraArgs ia, BondArgs ba, int nib) {
    __shared__ IntraLds L;
    if ((int)blockIdx.x < nib) intra_items<ENERGY>(ia, blockIdx.x, L);
    else bonded_lanes<ENERGY>(ba, blockIdx.x - nib);
}

// Sums the per-wave energy records of all regions in a fixed order, in two stages: block b adds
// the records of its contiguous share ([b n / B, (b+1) n / B)) into part[b], then one block adds the
// parts in block order into acc.  Records in [0, n1) are always present, [n1, n2) (kinetic) only
// when kin_valid.  The PME reciprocal energy comes from the per-slot 1/2 q phi array (ephi, when
// pme_valid), summed the same way.
__global__ __launch_bounds__(256) void k_reduce_energy_part(const double* __restrict__ epart, int n1, int n2, int kin_valid,
                                                            const float* __restrict__ ephi, int nslots, int pme_valid,
                                                            double* __restrict__ part) {
    __shared__ double red[ACC_N][4];
    double s[ACC_N];
    for (int c = 0; c < ACC_N; c++) s[c] = 0.0;
    const int nrec = kin_valid ? n2 : n1;
    const int r0 = (int)((int64_t)nrec * blockIdx.x / gridDim.x), r1 = (int)((int64_t)nrec * (blockIdx.x + 1) / gridDim.x);
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x)
        for (int c = 0; c < ACC_N; c++) s[c] += epart[(int64_t)r * ACC_N + c];
    if (pme_valid) {
        const int q0 = (int)((int64_t)nslots * blockIdx.x / gridDim.x), q1 = (int)((int64_t)nslots * (blockIdx.x + 1) / gridDim.x);
        for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) s[ACC_PME] += (double)ephi[q];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = 0; c < ACC_N; c++) {
        const double v = wave_sum_d(s[c]);
        if (lane == 0) red[c][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < ACC_N)
        part[blockIdx.x * ACC_N + threadIdx.x] = ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
}

__global__ void k_reduce_energy_final(const double* __restrict__ part, int nparts, double* __restrict__ acc) {
    const int c = threadIdx.x;
    if (c >= ACC_N) return;
    double t = 0.0;
    for (int b = 0; b < nparts; b++) t += part[b * ACC_N + c];
    acc[c] = t;
}

}  // namespace

int pair_split(const Engine& e) {
    // small boxes: split clusters until there are ~4 waves per SIMD (256 CUs x 4 SIMDs); a cluster's
    // waves stay inside one workgroup (split divides PAIR_WAVES).  Past that, one wave per cluster:
    // at 42k atoms all pair workgroups are then resident at once, and the slots they leave free take
    // the PME and intra / bonded kernels that run beside them (splitting further kept queueing pair
    // workgroups into every freed slot: the step was 3 % slower with the pair kernel's own time equal).
    int split = 1;
    while (split < PAIR_WAVES && (int64_t)e.nl.ncl_max * split < 4096) split *= 2;
    return split;
}

ERegions energy_regions(const Engine& e) {
    ERegions r;
    r.pair = 0;
    r.intra = r.pair + e.nl.ncl_max * pair_split(e);
    r.bond = r.intra + e.n_intra_groups;
    r.pme = r.bond + e.bd.n_lanes / 64;   // one record per wave of the bonded incidence lanes
    r.kin = r.pme + pme_virial_records(e);   // PME: the z pass's virial records (its energy comes from ephi)
    r.total = r.kin + (int)((e.N + 63) / 64);
    return r;
}

void launch_reduce_energy(Engine& e, bool kinetic, hipStream_t s) {
    const bool pme_valid = e.elec && e.pme.ncharged > 0;
    hipLaunchKernelGGL(k_reduce_energy_part, dim3(REDUCE_BLOCKS), dim3(256), 0, s, e.epart, e.ereg.kin, e.ereg.total,
                       kinetic ? 1 : 0, (const float*)e.ephi, slots_max(e), pme_valid ? 1 : 0, e.acc_part);
    hipLaunchKernelGGL(k_reduce_energy_final, dim3(1), dim3(64), 0, s, (const double*)e.acc_part, REDUCE_BLOCKS, e.acc);
}

void launch_pack(Engine& e, const uint4* w, hipStream_t s) {
    const int nslots = slots_max(e);
    hipLaunchKernelGGL(k_pack, dim3(nblk(nslots, 256)), dim3(256), 0, s, e.nl.ncl_dev, nslots, e.nl.catoms, w, e.p.prm, e.nl.cref,
                       (long long)e.Lq[0], (long long)e.Lq[1], (long long)e.Lq[2], e.nl.xq, bin_args(e));
}

void launch_pair(Engine& e, bool energy, hipStream_t s) {
    PairArgs a{};
    a.ncl_dev = e.nl.ncl_dev;
    a.ncl_max = e.nl.ncl_max;
    a.rc2 = e.rc2;
    a.zscale = (float)COUL_TAB / e.rc2;
    a.pc = e.pc;
    a.coul_tab = e.coul_tab;
    a.cap = e.nl.cap;
    a.xq = e.nl.xq; a.ljs = e.nl.ljs; a.ecount = e.nl.ecount; a.efar = e.nl.efar; a.ent = e.nl.ent;
    a.cref = e.nl.cref;
    for (int d = 0; d < 3; d++) a.lq[d] = (uint32_t)e.Lq[d];
    a.fpair = e.fpair; a.epart = e.epart + (int64_t)e.ereg.pair * ACC_N;
    a.split = pair_split(e);
    const int64_t waves = (int64_t)a.ncl_max * a.split;
    const dim3 grid(nblk(waves, PAIR_WAVES)), block(64 * PAIR_WAVES);
    if (energy) hipLaunchKernelGGL(k_pair<true>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(k_pair<false>, grid, block, 0, s, a);
}

IntraArgs intra_args(Engine& e, const uint4* w, const int4* img) {
    IntraArgs a{};
    a.n_groups = e.n_intra_groups;
    for (int d = 0; d < 3; d++) a.Lq[d] = e.Lq[d];
    a.pc = e.pc;
    a.groups = e.intra_groups; a.items = e.intra_items; a.lut = e.intra_lut; a.w = w; a.img = img; a.prm = e.p.prm;
    a.slot_of = e.nl.slot_of; a.first = e.first; a.fintra = e.fintra;
    a.epart = e.epart + (int64_t)e.ereg.intra * ACC_N;
    return a;
}

BondArgs bond_args(Engine& e, const uint4* w, const int4* img) {
    BondArgs a{};
    a.nb = e.bd.nb; a.na = e.bd.na; a.nd = e.bd.nd; a.ni = e.bd.ni; a.np = e.bd.np;
    for (int d = 0; d < 3; d++) a.Lq[d] = e.Lq[d];
    a.pc = e.pc;
    a.bid = e.bd.bid; a.bp = e.bd.bp; a.aid = e.bd.aid; a.ap = e.bd.ap; a.did = e.bd.did; a.dp = e.bd.dp;
    a.iid = e.bd.iid; a.ip = e.bd.ip; a.pid = e.bd.pid; a.pp = e.bd.pp;
    a.n_lanes = e.bd.n_lanes; a.multi_lane0 = e.bd.multi_lane0; a.lane_code = e.bd.lane_code; a.lane_atom = e.bd.lane_atom;
    a.w = w; a.img = img; a.slot_of = e.nl.slot_of; a.fb = e.fb;
    a.epart = e.epart + (int64_t)e.ereg.bond * ACC_N;
    return a;
}

void launch_local_blocks(Engine& e, const uint4* w, const int4* img, bool energy, bool intra, bool bonded, hipStream_t s) {
    const IntraArgs ia = intra_args(e, w, img);
    const BondArgs ba = bond_args(e, w, img);
    const int nib = intra ? nblk(ia.n_groups, 4) : 0;
    const int nbb = bonded ? nblk(ba.n_lanes, 256) : 0;
    if (nib + nbb == 0) return;
    if (energy) hipLaunchKernelGGL(k_local<true>, dim3(nib + nbb), dim3(256), 0, s, ia, ba, nib);
    else hipLaunchKernelGGL(k_local<false>, dim3(nib + nbb), dim3(256), 0, s, ia, ba, nib);
}

void launch_local(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s) {
    launch_local_blocks(e, w, img, energy, true, true, s);
}

void launch_intra(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s) {
    launch_local_blocks(e, w, img, energy, true, false, s);
}

void launch_bonded(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s) {
    launch_local_blocks(e, w, img, energy, false, true, s);
}

}  // namespace lima
