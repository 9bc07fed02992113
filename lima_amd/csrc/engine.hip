// engine.hip -- C ABI of liblima_hip.so (include/lima_hip.h): engine lifetime, MD step,
// steepest-descent minimisation, state read-back.
//
// Stands in for LIMA_ENGINE's Engine (code/LIMA_ENGINE/include/Engine.cuh:78-170,
// Engine.cu:19-463) and Programs::EnergyMinimize (code/LIMA_MD/src/Programs.cpp:155-199).
// Positions are integer Coords (1e-8 nm, LimaTypes.cuh:297-334) kept per particle as a wrapped
// uint32 plus an image counter, so the unwrapped coordinate LIMA's compound-relative layout
// represents is recoverable exactly and periodic wrapping never loses bits.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "lima_terms.h"

using namespace lima;

struct lima_engine {
    Engine e;
};

namespace lima {
thread_local std::string g_err;   // message of the last failing C-ABI call on this thread
}

namespace {

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

int64_t fdiv(int64_t a, int64_t b) { int64_t q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) --q; return q; }

// ---- integrator, EngineUtils.cuh:31-41 / EngineKernels.cuh:383-397 ----
// v(t) = v(t-dt) + (F(t) + F(t-dt)) dt/2m ;  x += Coord{ v(t) dt + F(t) dt^2 / 2m } (truncating)
struct IntegrateArgs {
    int64_t N;
    long long Lq[3];
    float dt;
    const float* thermostat;   // [0] velocity scale factor (Thermostat.cuh:56-67), set on the device
    const uint4* w_build;
    const int4* img_build;
    float2* dpart;         // per wave: running maxima of the two largest squared displacements since the list build
    double* epart;         // kinetic-energy region: one record per wave
    const int* slot_of;
    const float4* fpair;   // per-kernel force records [nslots], summed in this order
    const float4* fintra;
    const float4* fb;
    const float4* fpme;
    uint4* w;
    int4* img;
    float4* v;
    float4* fprev;
    const float4* prm;
    // pack != 0: also prepare the particle's slot for the next evaluation, as k_pack does (relative
    // position and charge, PME tile bin); the list must stay unchanged
    int pack;
    const uint4* cref;
    float4* xq;
    BinArgs bins;
    int* flags;            // DF_NONFINITE is set when a force is NaN / inf
};

// w + d wrapped back into [0, Lq) for any d (a step may exceed a box edge when a run explodes).
__device__ __forceinline__ void wrap_add(uint32_t& w, int& img, long long d, long long Lq) {
    long long x = (long long)w + d;
    if (x >= Lq || x < 0) {
        long long q = x / Lq;
        if (x - q * Lq < 0) q--;
        x -= q * Lq;
        img += (int)q;
    }
    w = (uint32_t)x;
}

__device__ __forceinline__ int coord_of(float nm) {   // Coord{Float3}: static_cast<int32_t>(pos * 1e8)
    const float v = nm * LIMA_NANO_TO_LIMA_F;
    return (int)fminf(fmaxf(v, -2.0e9f), 2.0e9f);
}

template <bool ENERGY>
__global__ void k_integrate(IntegrateArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t wv = p >> 6;
    const bool lead = (threadIdx.x & 63) == 0 && (p & ~63ll) < a.N;   // lane 0 of a wave inside the per-wave arrays
    // the wave's running displacement maxima, loaded first: independent of every other load
    const float2 dp_old = lead ? a.dpart[wv] : make_float2(0.f, 0.f);
    float ke = 0.f, d2 = 0.f;
    uint4 bin_w = make_uint4(0u, 0u, 0u, 0u);
    bool bin_on = false;   // pack: this particle goes to a PME tile bin
    int bin_s = 0;
    float bin_q = 0.f;
    bool bad = false;      // non-finite force
    if (p < a.N) {
        const int s = a.slot_of[p];
        // every load before the first store: the compiler cannot tell the arrays apart, and loads
        // issued after a store would each add a memory round trip to this latency-bound kernel
        const float4 f0 = a.fpair[s], f1 = a.fintra[s], f2 = a.fb[s], f3 = a.fpme[s];
        const float fx = ((f0.x + f1.x) + f2.x) + f3.x;
        const float fy = ((f0.y + f1.y) + f2.y) + f3.y;
        const float fz = ((f0.z + f1.z) + f2.z) + f3.z;
        bad = !(fabsf(fx) <= 3.0e38f && fabsf(fy) <= 3.0e38f && fabsf(fz) <= 3.0e38f);
        const float4 prm = a.prm[p];
        const float m = prm.w;
        const float4 fp = a.fprev[p];
        float4 v = a.v[p];
        uint4 w = a.w[p];
        int4 im = a.img[p];
        const uint4 wb = a.w_build[p];
        const int4 ib = a.img_build[p];
        const uint4 r = a.pack ? a.cref[s >> 3] : make_uint4(0u, 0u, 0u, 0u);
        const float ts = a.thermostat[0];
        const float kdt = a.dt * 0.5f / m;
        v.x = v.x + (fx + fp.x) * kdt;
        v.y = v.y + (fy + fp.y) * kdt;
        v.z = v.z + (fz + fp.z) * kdt;
        const float c2 = 0.5f / m * a.dt * a.dt;
        wrap_add(w.x, im.x, coord_of(v.x * a.dt + fx * c2), a.Lq[0]);
        wrap_add(w.y, im.y, coord_of(v.y * a.dt + fy * c2), a.Lq[1]);
        wrap_add(w.z, im.z, coord_of(v.z * a.dt + fz * c2), a.Lq[2]);
        a.w[p] = w;
        a.img[p] = im;
        if (a.pack) {
            long long d[3] = {(long long)w.x - (long long)r.x, (long long)w.y - (long long)r.y, (long long)w.z - (long long)r.z};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (2 * d[k] > a.Lq[k]) d[k] -= a.Lq[k];
                else if (2 * d[k] < -a.Lq[k]) d[k] += a.Lq[k];
            }
            const float q = prm.x;
            a.xq[s] = make_float4((float)d[0] * LIMA_LIMA_TO_NANO_F, (float)d[1] * LIMA_LIMA_TO_NANO_F, (float)d[2] * LIMA_LIMA_TO_NANO_F, q);
            bin_w = w;
            bin_on = q != 0.f;
            bin_s = s;
            bin_q = q;
        }
        {   // displacement since the list was built (unwrapped, exact integer difference)
            const float dx = (float)((long long)w.x - (long long)wb.x + (long long)(im.x - ib.x) * a.Lq[0]) * LIMA_LIMA_TO_NANO_F;
            const float dy = (float)((long long)w.y - (long long)wb.y + (long long)(im.y - ib.y) * a.Lq[1]) * LIMA_LIMA_TO_NANO_F;
            const float dz = (float)((long long)w.z - (long long)wb.z + (long long)(im.z - ib.z) * a.Lq[2]) * LIMA_LIMA_TO_NANO_F;
            d2 = dx * dx + dy * dy + dz * dz;
        }
        if (ENERGY) ke = 0.5f * m * (v.x * v.x + v.y * v.y + v.z * v.z);
        v.x *= ts; v.y *= ts; v.z *= ts;
        a.v[p] = v;
        a.fprev[p] = make_float4(fx, fy, fz, 0.f);
    }
    if (a.pack && a.bins.bin_count) bin_wave(a.bins, bin_on, bin_s, bin_w, bin_q);   // all lanes, converged
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&a.flags[DF_NONFINITE], 1);
    // the wave's largest and second-largest squared displacement (lanes past N hold 0)
    const float m1 = wave_max(d2);
    const bool top = d2 == m1;
    const float m2 = __popcll(__ballot(top)) > 1 ? m1 : wave_max(top ? 0.f : d2);
    if (lead) a.dpart[wv] = make_float2(fmaxf(dp_old.x, m1), fmaxf(dp_old.y, m2));
    if (ENERGY) {
        const double t = wave_sum_d((double)ke);
        if (lead) {
            double* r = a.epart + wv * ACC_N;
            write_record(r, 0.0, 0.0, 0.0, t, 0.0, nullptr);
        }
    }
}

__global__ void k_kinetic(int64_t N, const float4* __restrict__ v, const float4* __restrict__ prm, double* epart) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float ke = 0.f;
    if (p < N) {
        const float4 u = v[p];
        ke = 0.5f * prm[p].w * (u.x * u.x + u.y * u.y + u.z * u.z);
    }
    const double t = wave_sum_d((double)ke);
    if ((threadIdx.x & 63) == 0 && (p & ~63ll) < N) {
        double* r = epart + (p >> 6) * ACC_N;
        write_record(r, 0.0, 0.0, 0.0, t, 0.0, nullptr);
    }
}

// Thermostat::Temperature (Thermostat.cuh:87-103) with ComputeThermostatScalar (:56-67): the
// kinetic energy of the stored velocities is summed in FP64 in a fixed order, T = 2 KE / (dof k_B N_A)
// with dof = 3N - 3 (BoxBuilder.cpp:108), and the velocity scale is clamp(310 K / T, 1 -+ max_scalar),
// max_scalar = 0.001 / steps_per_temperature_measurement.  The temperature is always recorded; the
// scale the integrator applies from the next step on is replaced only when the thermostat is on
// (Engine.cu:167-174).
__global__ __launch_bounds__(256) void k_thermostat(const double* __restrict__ kin, int nrec, double dof,
                                                    float max_scalar, int apply, float* thermo) {
    __shared__ double red[256];
    double t = 0.0;
    for (int r = threadIdx.x; r < nrec; r += 256) t += kin[(int64_t)r * ACC_N + ACC_KIN];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float T = dof > 0.0 ? (float)(red[0] * (2.0 / dof) / ((double)1.38066e-23f * 6.02214076e23)) : 0.f;
        const float safe = T == 0.f ? 1.f : T;
        const float sc = fminf(fmaxf(310.f / safe, 1.f - max_scalar), 1.f + max_scalar);
        thermo[1] = T;
        if (apply) thermo[0] = sc;
    }
}

// The per-kernel force and potE buffers of the last evaluation, summed in the integrator's order.
struct SlotForces {
    const float4* f[4];   // pair, intra, bonded, PME: [nslots] (fx, fy, fz, potE share)
};
__device__ __forceinline__ float4 slot_force(const SlotForces& b, int s) {
    const float4 f0 = b.f[0][s], f1 = b.f[1][s], f2 = b.f[2][s], f3 = b.f[3][s];
    return make_float4(((f0.x + f1.x) + f2.x) + f3.x, ((f0.y + f1.y) + f2.y) + f3.y, ((f0.z + f1.z) + f2.z) + f3.z,
                       ((f0.w + f1.w) + f2.w) + f3.w);
}

// Forces of the last evaluation in global order (w = per-particle potE) and max |F|.
__global__ void k_forces_global(int64_t N, const int* __restrict__ slot_of, SlotForces b, float4* __restrict__ out, float* fmax) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.f;
    if (p < N) {
        const int s = slot_of[p];
        const float4 f = slot_force(b, s);
        out[p] = f;
        m = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z);
        if (!(m <= 3.0e38f)) m = 3.0e38f;   // NaN / inf -> flagged as huge
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_pos(fmax, m);
}

// One data-log frame (LogCompoundData / LogSolventData, EngineUtils.cuh:143-173): position after
// the integration in LIMA's convention (compound particles: origo = the key particle's wrapped nm
// node, Coord relative to it, so a compound stays whole; tinymols: wrapped into the box), potE and
// force of the step's evaluation, speed after the step.  Written to the particle's entry of LIMA's
// buffer layout.
__global__ void k_log_frame(int64_t N, const int* __restrict__ lidx, const int* __restrict__ key, const uint4* __restrict__ w,
                            const int4* __restrict__ img, const float4* __restrict__ v, const int* __restrict__ slot_of, SlotForces b,
                            long long Lqx, long long Lqy, long long Lqz, float4* __restrict__ pe, float4* __restrict__ fv) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const uint4 wp = w[p];
    const int k = key[p];
    const long long wpd[3] = {(long long)wp.x, (long long)wp.y, (long long)wp.z};
    long long node[3], coord[3];
    if (k >= 0) {
        const uint4 wk = w[k];
        const int4 ik = img[k], ip = img[p];
        const long long wkd[3] = {(long long)wk.x, (long long)wk.y, (long long)wk.z};
        const long long dimg[3] = {(long long)(ip.x - ik.x), (long long)(ip.y - ik.y), (long long)(ip.z - ik.z)};
        const long long Lq[3] = {Lqx, Lqy, Lqz};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            node[d] = wkd[d] / LIMA_NANO_TO_LIMA;
            coord[d] = wpd[d] - wkd[d] + dimg[d] * Lq[d] + (wkd[d] - node[d] * LIMA_NANO_TO_LIMA);
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            node[d] = wpd[d] / LIMA_NANO_TO_LIMA;
            coord[d] = wpd[d] - node[d] * LIMA_NANO_TO_LIMA;
        }
    }
    // FP32 origo + Coord * 1e-8 and |v| (the multiply-adds contract to FMAs, as in LIMA's build)
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; d++) x[d] = (float)node[d] + (float)(int)coord[d] * LIMA_LIMA_TO_NANO_F;
    const int s = slot_of[p];
    const float4 vp = v[p];
    const float speed = sqrtf(vp.x * vp.x + vp.y * vp.y + vp.z * vp.z);
    const int64_t o = lidx[p];
    const float4 f = slot_force(b, s);
    pe[o] = make_float4(x[0], x[1], x[2], f.w);
    fv[o] = make_float4(f.x, f.y, f.z, speed);
}

__global__ void k_zero_d(double* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}
__global__ void k_zero_f(float* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.f;
}

// ---- steepest descent (GROMACS integrator=steep, dev/gmxscripts/em.mdp:2-5) ----
__global__ void k_sd_trial(int64_t N, long long Lqx, long long Lqy, long long Lqz, const SdState* sd,
                           const uint4* __restrict__ w, const int4* __restrict__ img, const float4* __restrict__ f,
                           uint4* __restrict__ wt, int4* __restrict__ it) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const float k = sd->h / sd->fmax_acc;
    const float4 F = f[p];
    uint4 x = w[p];
    int4 m = img[p];
    wrap_add(x.x, m.x, coord_of(k * F.x), Lqx);
    wrap_add(x.y, m.y, coord_of(k * F.y), Lqy);
    wrap_add(x.z, m.z, coord_of(k * F.z), Lqz);
    wt[p] = x;
    it[p] = m;
}

// flags[DF_PMEBIN] set: the trial's PME forces are incomplete, so it is refused and h kept (the
// host grows the bins and repeats the iteration).
__global__ void k_sd_accept(SdState* sd, const double* acc, double e_const, const float* fmax_try, const int* flags) {
    const double e = acc[ACC_LJ] + acc[ACC_COUL] + acc[ACC_BOND] + acc[ACC_PME] + e_const;
    sd->e_try = e;
    sd->fmax_try = *fmax_try;
    if (flags[DF_PMEBIN]) {
        sd->accepted = 0;
    } else if (e < sd->e_acc) {
        sd->accepted = 1;
        sd->disp_since_build += sd->h;
        sd->e_acc = e;
        sd->fmax_acc = *fmax_try;
        sd->h *= 1.2f;
        sd->n_accepted++;
    } else {
        sd->accepted = 0;
        sd->h *= 0.2f;
    }
}

__global__ void k_sd_commit(int64_t N, const SdState* sd, const uint4* __restrict__ wt, const int4* __restrict__ it,
                            const float4* __restrict__ ft, uint4* __restrict__ w, int4* __restrict__ img, float4* __restrict__ facc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N || !sd->accepted) return;
    w[p] = wt[p];
    img[p] = it[p];
    facc[p] = ft[p];
}

// ---------------------------------------------------------------------------------------------

void set_err(const std::string& s) { g_err = s; }

template <class F>
lima_status guarded(F&& f) {
    try {
        f();
        return LIMA_OK;
    } catch (const LimaFail& x) {
        set_err(x.what());
        return x.code;
    } catch (const HipError& x) {
        set_err(x.what());
        return LIMA_ERR_HIP;
    } catch (const std::exception& x) {
        set_err(x.what());
        return LIMA_ERR_BAD_INPUT;
    }
}

template <class T>
T* dalloc_copy(const std::vector<T>& v) {
    T* p = nullptr;
    if (v.empty()) { HIP_CHECK(hipMalloc(&p, sizeof(T))); return p; }
    HIP_CHECK(hipMalloc(&p, v.size() * sizeof(T)));
    HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

double e_const(const Engine& e) { return e.elec ? e.pme.eself_total : 0.0; }

hipEvent_t* tslot(Engine& e) { return &e.tev[e.tpos * Engine::TEV]; }

void timing_flush(Engine& e) {
    if (e.tpos == 0) return;
    HIP_CHECK(hipEventSynchronize(e.tev[(e.tpos - 1) * Engine::TEV + Engine::TEV - 1]));
    for (int k = 0; k < e.tpos; k++) {
        hipEvent_t* ev = &e.tev[k * Engine::TEV];
        float t[6];
        for (int i = 0; i < 6; i++) HIP_CHECK(hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
        e.tacc.pair_ms += t[1];
        e.tacc.intra_ms += t[2];
        e.tacc.bonded_ms += t[3];
        e.tacc.pme_ms += t[4];
        e.tacc.integrate_ms += t[0] + t[5];   // pack + integrate: the per-particle streaming kernels
        float tot;
        HIP_CHECK(hipEventElapsedTime(&tot, ev[0], ev[6]));
        e.tacc.total_ms += tot;
        e.tacc.steps_timed++;
    }
    e.tpos = 0;
}

// One force evaluation.  The intra / bonded kernel and PME (spread, FFT, gather) run on a second
// stream between a fork after the pack kernel (which zeroes the force slots) and a join before
// the caller's next kernel, overlapping the pair kernel (LIMA overlaps PME the same way,
// Engine.cu:333).  With timing on, everything runs on one stream so each phase is timed alone.
// pack = false: the slots were already packed for these positions (by the integrator of the
// previous step, or by an explicit k_pack before a replayed step graph).
void evaluate(Engine& e, const uint4* w, const int4* img, bool energy, bool timed = false, bool pack = true) {
    hipStream_t s = e.s_main;
    hipEvent_t* ev = timed ? tslot(e) : nullptr;
    if (ev) HIP_CHECK(hipEventRecord(ev[0], s));
    if (pack) launch_pack(e, w, s);
    e.packed = false;
    if (ev) HIP_CHECK(hipEventRecord(ev[1], s));
    // Two streams after the pack: the intra-compound pairs and bonded terms (k_local,
    // latency-bound), then PME, on s_pme beside the VALU-bound pair kernel on the main stream.
    // They write disjoint force and energy buffers; the main stream joins s_pme before the
    // integrator.  The side branch is captured first: a replayed graph puts the first branch
    // after a fork on the launching queue, and a third branch costs one more cross-queue join.
    const bool split = !timed && !(e.prm.launch_mode & LAUNCH_ONE_STREAM);
    const bool pme_on = e.elec && e.pme.ncharged > 0;
    if (split) {
        HIP_CHECK(hipEventRecord(e.ev_fork, s));
        HIP_CHECK(hipStreamWaitEvent(e.s_pme, e.ev_fork, 0));
        launch_local(e, w, img, energy, e.s_pme);
        if (pme_on) launch_pme(e, w, energy, e.s_pme);
        HIP_CHECK(hipEventRecord(e.ev_join, e.s_pme));
    }
    launch_pair(e, energy, s);
    if (ev) HIP_CHECK(hipEventRecord(ev[2], s));
    if (ev) {   // timed: each phase alone
        launch_intra(e, w, img, energy, s);
        HIP_CHECK(hipEventRecord(ev[3], s));
        launch_bonded(e, w, img, energy, s);
        HIP_CHECK(hipEventRecord(ev[4], s));
    } else if (!split) {
        launch_local(e, w, img, energy, s);
    }
    if (split) {
        HIP_CHECK(hipStreamWaitEvent(s, e.ev_join, 0));
    } else {
        launch_pme(e, w, energy, s);
    }
    if (ev) HIP_CHECK(hipEventRecord(ev[5], s));
}

// lima_energy from the FP64 accumulators of one evaluation (E.kinetic kept when acc has none).
void energy_from_acc(const Engine& e, const double* acc, bool with_kinetic, lima_energy& E) {
    E.lj = acc[ACC_LJ];
    E.coulomb_sr = acc[ACC_COUL];
    E.bonded = acc[ACC_BOND];
    E.pme_recip = acc[ACC_PME];
    E.pme_self_ewald = e.elec ? e.pme.eself_total : 0.0;
    E.pme_self_lima = e.elec ? 0.5 * e.pme.eself_total * (double)e.pme.ncharged : 0.0;
    E.potential = E.lj + E.coulomb_sr + E.bonded + E.pme_recip + E.pme_self_ewald;
    E.potential_lima = E.lj + E.coulomb_sr + E.bonded + E.pme_recip + E.pme_self_lima;
    if (with_kinetic) E.kinetic = acc[ACC_KIN];
    E.total = E.potential + E.kinetic;
    for (int k = 0; k < 6; k++) E.virial[k] = acc[ACC_VXX + k];
}

// Reads the totals of the last energy evaluation (e.acc) into e.energy; the device must be idle.
void fill_energy(Engine& e, bool with_kinetic, int64_t step) {
    double acc[ACC_N];
    HIP_CHECK(hipMemcpy(acc, e.acc, sizeof(acc), hipMemcpyDeviceToHost));
    energy_from_acc(e, acc, with_kinetic, e.energy);
    e.energy.step = step;
    e.pairs_last = (int64_t)acc[ACC_PAIRS];
    e.energy_unread = false;
}

SlotForces slot_forces(const Engine& e) {
    SlotForces b{};
    b.f[0] = e.fpair; b.f[1] = e.fintra; b.f[2] = e.fb; b.f[3] = e.fpme;
    return b;
}

// Writes the current step's frame into the data-log ring (after the integration and the energy
// reduction of a logged step, on the main stream).
void log_frame(Engine& e) {
    const int64_t slot = e.log_frames % e.log_cap;
    hipLaunchKernelGGL(k_log_frame, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, e.N, e.log_index, e.log_key, e.p.w, e.p.img,
                       e.p.v, e.nl.slot_of, slot_forces(e), (long long)e.Lq[0], (long long)e.Lq[1],
                       (long long)e.Lq[2], e.log_pe + slot * e.log_P, e.log_fv + slot * e.log_P);
    HIP_CHECK(hipMemcpyAsync(e.log_acc + slot * ACC_N, e.acc, ACC_N * sizeof(double), hipMemcpyDeviceToDevice, e.s_main));
    e.log_step[slot] = e.step;
    e.log_frames++;
}

void check_finite(Engine& e) {
    if (!std::isfinite(e.energy.potential)) {
        e.critical = true;
        throw LimaFail(LIMA_ERR_NONFINITE, "non-finite energy");
    }
}

// pack: the next evaluation uses the same list, so the integrator packs the slots for it and
// the evaluation skips k_pack (one launch fewer per step).
void integrate(Engine& e, bool energy, bool pack = false) {
    IntegrateArgs a{};
    a.N = e.N;
    for (int d = 0; d < 3; d++) a.Lq[d] = e.Lq[d];
    a.dt = (float)e.dt_ns;
    a.thermostat = e.thermo;
    a.slot_of = e.nl.slot_of;
    a.fpair = e.fpair; a.fintra = e.fintra; a.fb = e.fb; a.fpme = e.fpme;
    a.w = e.p.w; a.img = e.p.img; a.v = e.p.v; a.fprev = e.p.fprev;
    a.prm = e.p.prm; a.epart = e.epart + (int64_t)e.ereg.kin * ACC_N;
    a.w_build = e.w_build; a.img_build = e.img_build; a.dpart = e.dpart;
    a.pack = pack ? 1 : 0;
    a.cref = e.nl.cref; a.xq = e.nl.xq;
    if (pack) a.bins = bin_args(e);
    a.flags = e.dflags;
    e.packed = pack;
    if (energy) hipLaunchKernelGGL(k_integrate<true>, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, a);
    else hipLaunchKernelGGL(k_integrate<false>, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, a);
}

// LIMA measures the temperature in hostMaster when the logging buffer is full and the step is a
// multiple of steps_per_temperature_measurement (Engine.cu:162-174; IsBufferFull is
// step % (5 * data_logging_interval) == 0, SimulationDevice.cuh:50-54).
bool temperature_due(const Engine& e) {
    const int64_t spt = e.prm.steps_per_temperature_measurement;
    const int64_t buf = 5 * (int64_t)e.prm.data_logging_interval;
    return spt > 0 && buf > 0 && e.step > 0 && e.step % buf == 0 && e.step % spt == 0;
}

void measure_temperature(Engine& e) {
    double* kin = e.epart + (int64_t)e.ereg.kin * ACC_N;
    hipLaunchKernelGGL(k_kinetic, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, e.N, e.p.v, e.p.prm, kin);
    const float max_scalar = 0.001f / (float)e.prm.steps_per_temperature_measurement;
    hipLaunchKernelGGL(k_thermostat, dim3(1), dim3(256), 0, e.s_main, (const double*)kin, (int)((e.N + 63) / 64),
                       3.0 * (double)e.N - 3.0, max_scalar, e.prm.apply_thermostat ? 1 : 0, e.thermo);
    e.temp_measured = true;
}

// Captures one MD step without energies (local terms and PME on the second stream, pair, join,
// integrator and the next step's pack) into a graph.  Replaying it replaces ~16 host launch and event calls per step
// by one graph launch; the kernels, their arguments and the fork/join are those of the eager path.
// An existing executable graph is updated in place when the topology is unchanged.
void capture_step_graph(Engine& e) {
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(e.s_main, hipStreamCaptureModeThreadLocal));
    try {
        evaluate(e, e.p.w, e.p.img, false, false, /*pack=*/false);
        integrate(e, false, /*pack=*/true);
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(e.s_main, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(e.s_main, &g));
    if (e.step_graph) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(e.step_graph, g, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(e.step_graph);
            e.step_graph = nullptr;
        }
    }
    if (!e.step_graph) {
        const hipError_t err = hipGraphInstantiate(&e.step_graph, g, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            (void)hipGraphDestroy(g);
            e.step_graph = nullptr;
            throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
        }
    }
    HIP_CHECK(hipGraphDestroy(g));
    e.step_graph_valid = true;
    e.graph_captures++;
}

// Undoes the steps since the last snapshot after a device condition: restores the snapshot's
// state, grows the capacity that overflowed, and leaves e.step at the snapshot's step with a
// rebuild due, so the caller issues the lost steps again.  A non-finite force or an overfull
// column is not recoverable and throws.
void rollback(Engine& e) {
    HIP_CHECK(hipStreamSynchronize(e.s_pme));
    HIP_CHECK(hipStreamSynchronize(e.s_main));
    int f[DF_N];
    HIP_CHECK(hipMemcpy(f, e.dflags, sizeof(f), hipMemcpyDeviceToHost));
    e.rb_pending[0] = e.rb_pending[1] = false;
    if (f[DF_NONFINITE]) {
        e.critical = true;
        throw LimaFail(LIMA_ERR_NONFINITE, "non-finite force in the MD step loop");
    }
    if (f[DF_COLUMN]) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "an x-y column holds more than 2048 particles");
    const int64_t c = f[DF_SNAPSHOTS];
    auto it = std::find_if(e.snaps.begin(), e.snaps.end(), [&](const SnapRecord& r) { return r.ordinal == c; });
    if (it == e.snaps.end()) throw LimaFail(LIMA_ERR_HIP, "internal: no record of the last snapshot");
    const SnapRecord r = *it;
    if (f[DF_PAIRLIST] || f[DF_PMEBIN]) {
        if (++e.rollbacks > 64) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "capacity overflows persist after 64 rollbacks");
        if (f[DF_PAIRLIST]) nlist_grow(e, f[DF_LIST_NEED]);
        if (f[DF_PMEBIN]) pme_grow_bins(e);
    }
    e.list_interval = r.list_interval;   // the schedule in force when the snapshot was taken
    e.builds_clean = r.builds_clean;
    if (f[DF_STALE]) {   // run the interval again with lists that live a sixth shorter
        if (e.list_interval <= 1)
            throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "particles move more than half the list skin in one step");
        e.list_interval -= std::max(1, e.list_interval / 6);
        e.builds_clean = 0;
        e.grow_after = std::min(2 * e.grow_after, 640);
        e.stale_rollbacks++;
    }
    hipStream_t s = e.s_main;
    const int64_t N = e.N;
    HIP_CHECK(hipMemcpyAsync(e.p.w, e.w_build, N * sizeof(uint4), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.p.img, e.img_build, N * sizeof(int4), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.p.v, e.v_snap, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.p.fprev, e.fprev_snap, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.thermo, e.thermo_snap, 2 * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.lstats, e.lstats_snap, 2 * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(e.acc, e.acc_snap, ACC_N * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemsetAsync(e.dpart, 0, ((N + 63) / 64) * sizeof(float2), s));
    pme_clear_bins(e, s);
    reset_flags(e, c - 1, s);
    HIP_CHECK(hipStreamSynchronize(s));
    e.snap_issued = c - 1;
    e.snaps.erase(it, e.snaps.end());
    e.step = r.step;
    e.log_frames = r.log_frames;
    e.n_rebuilds = r.n_rebuilds;
    e.temp_measured = r.temp_measured;
    e.energy_step = r.energy_step;
    e.energy_unread = r.energy_step != -1;   // acc holds that evaluation's totals again
    e.steps_since_build = INT64_MAX / 2;      // the next step rebuilds from the restored positions
    e.stats_skip = true;
    e.packed = false;
    e.tpos = 0;                               // timing events of undone steps are dropped
}

// Reads the status words the device copied after the deferred rebuild issued two rebuilds ago,
// waiting for that rebuild if it is still running (this bounds how far the host runs ahead of the
// device).  Returns true when a condition was found and the engine rolled back.
bool poll_rebuild_flags(Engine& e) {
    const int par = e.rb_parity;
    if (!e.rb_pending[par]) return false;
    HIP_CHECK(hipEventSynchronize(e.ev_rebuild[par]));
    e.rb_pending[par] = false;
    const int* f = e.hflags + par * DF_N;
    bool err = false;
    for (int k = 0; k < DF_ERRORS; k++) err = err || f[k] != 0;
    if (!err) return false;
    rollback(e);
    return true;
}

// Issues MD steps until e.step == target, without waiting for the device (it is at most two list
// intervals behind, see poll_rebuild_flags).  Energies are evaluated on every energy_interval-th
// step and on logged steps; their totals stay on the device until read (settle).
void issue_steps(Engine& e, int64_t target) {
    const auto t_issue = std::chrono::steady_clock::now();
    const int64_t step0 = e.step;
    const int max_interval = e.prm.nlist_interval > 0 ? e.prm.nlist_interval : 20;
    while (e.step < target) {
        const int interval = e.list_interval;   // a rollback may shorten it
        if (e.steps_since_build >= interval) {
            if (poll_rebuild_flags(e)) continue;   // rolled back to an earlier step: issue from there
            // after grow_after rebuilds without a stale interval, try lists one step longer
            if (++e.builds_clean >= e.grow_after && e.list_interval < max_interval) {
                e.list_interval++;
                e.builds_clean = 0;
            }
            // timing: drain earlier steps first, so the rebuild's clock holds the rebuild alone
            if (e.timing) HIP_CHECK(hipDeviceSynchronize());
            const auto t0 = std::chrono::steady_clock::now();
            nlist_build(e, e.p.w, /*checked=*/false);
            if (e.timing) {
                HIP_CHECK(hipStreamSynchronize(e.s_main));
                e.tacc.nlist_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
        }
        const bool energy_step = e.prm.energy_interval > 0 && e.step % e.prm.energy_interval == 0;
        const bool logged = e.log_cap > 0 && e.step % e.prm.data_logging_interval == 0;
        const bool energy = energy_step || logged;
        if (!energy && !e.timing && !(e.prm.launch_mode & LAUNCH_EAGER)) {
            // the graph holds the evaluation without k_pack and an integrator that packs for the next
            // step; after a list rebuild (or an unpacked eager step) k_pack runs first
            if (!e.packed) launch_pack(e, e.p.w, e.s_main);
            if (!e.step_graph_valid) capture_step_graph(e);
            HIP_CHECK(hipGraphLaunch(e.step_graph, e.s_main));
            e.packed = true;
        } else {
            evaluate(e, e.p.w, e.p.img, energy, e.timing, !e.packed);
            // the integrator packs for the next step when that step keeps the list (the force
            // buffers are not touched by packing, so the last evaluation stays readable)
            const bool pack_next = !e.timing && e.steps_since_build + 1 < interval;
            integrate(e, energy, pack_next);
        }
        if (energy) {
            launch_reduce_energy(e, true, e.s_main);
            e.energy_step = e.step;
            e.energy_unread = true;
        }
        if (logged) log_frame(e);
        if (e.timing) {
            HIP_CHECK(hipEventRecord(tslot(e)[6], e.s_main));
            if (++e.tpos == Engine::TRING) timing_flush(e);
        }
        e.step++;
        e.steps_since_build++;
        if (temperature_due(e)) measure_temperature(e);
    }
    e.forces_valid = true;
    e.host_issue_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_issue).count();
    e.host_steps += std::max<int64_t>(0, e.step - step0);
    HIP_CHECK(hipGetLastError());
}

// Waits for every issued step, resolves device conditions (rolling back and issuing again as
// often as needed), and reads unread energy totals.  Every call that returns data runs it first.
void settle(Engine& e) {
    for (;;) {
        HIP_CHECK(hipStreamSynchronize(e.s_pme));
        HIP_CHECK(hipStreamSynchronize(e.s_main));
        HIP_CHECK(hipGetLastError());
        // the interval still open has had no rebuild to check it: its completeness bound now
        if (e.steps_since_build > 0 && e.steps_since_build < INT64_MAX / 4) list_interval_check(e, e.s_main);
        int f[DF_N];
        HIP_CHECK(hipMemcpy(f, e.dflags, sizeof(f), hipMemcpyDeviceToHost));
        e.rb_pending[0] = e.rb_pending[1] = false;
        bool err = false;
        for (int k = 0; k < DF_ERRORS; k++) err = err || f[k] != 0;
        if (!err) break;
        const int64_t target = e.step;
        rollback(e);
        issue_steps(e, target);
    }
    if (e.energy_unread) {
        fill_energy(e, true, e.energy_step);
        check_finite(e);
    }
}

void create(Engine& e, const lima_box_desc* b, const lima_simparams* p) {
    if (!b || !p) throw LimaFail(LIMA_ERR_BAD_INPUT, "null descriptor");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw LimaFail(LIMA_ERR_NO_DEVICE, "no HIP device visible");
    e.prm = *p;
    if (p->launch_mode < 0 || p->launch_mode > (LAUNCH_EAGER | LAUNCH_ONE_STREAM))
        throw LimaFail(LIMA_ERR_BAD_INPUT, "launch_mode must be a combination of LIMA_LAUNCH_EAGER and LIMA_LAUNCH_ONE_STREAM");
    if (p->apply_thermostat && p->steps_per_temperature_measurement <= 0)
        throw LimaFail(LIMA_ERR_BAD_INPUT, "the thermostat needs steps_per_temperature_measurement > 0");
    if (p->log_capacity < 0 || (p->log_capacity > 0 && p->data_logging_interval <= 0))
        throw LimaFail(LIMA_ERR_BAD_INPUT, "the data log needs log_capacity >= 0 and, when enabled, data_logging_interval > 0");
    if (p->bc_select != LIMA_BC_PBC) throw LimaFail(LIMA_ERR_UNSUPPORTED, "only periodic boundary conditions are implemented");
    if (!(p->cutoff_nm > 0.f)) throw LimaFail(LIMA_ERR_BAD_INPUT, "cutoff must be positive");
    e.nC = b->n_compounds;
    e.nT = b->n_tinymols;
    if (e.nC < 0 || e.nT < 0) throw LimaFail(LIMA_ERR_BAD_INPUT, "negative counts");
    e.rc = p->cutoff_nm;
    e.rc2 = e.rc * e.rc;
    e.skin = p->nlist_skin_nm > 0 ? p->nlist_skin_nm : 0.15f;
    // lists start short and grow while they stay complete (a stale interval costs a rerun)
    e.list_interval = std::min(p->nlist_interval > 0 ? p->nlist_interval : 40, 10);
    e.elec = p->enable_electrostatics != 0;
    e.pc.kc = (float)(8.9875517873681764e9 / 1e-9 / 6.02214076e23 * 1000.f * 1000.f);   // PhysicsUtils.cuh:5
    e.pc.kappa = 3.f / e.rc;                                                             // PhysicsUtils.cuh:60-62
    e.pc.two_kappa_sqrtpi = 2.f * e.pc.kappa / sqrtf(LIMA_PI_F);
    {   // H(r) = erfc(kappa r) + c r exp(-kappa^2 r^2), c = 2 kappa / sqrt(PI_L), and
        // H'(r) = (c (1 - 2 kappa^2 r^2) - 2 kappa / sqrt(pi)) exp(-kappa^2 r^2), in double, tabulated
        // in z = r^2 with dH/dz = H'(r) / (2 r) (its limit at r = 0 is 0)
        const double kap = (double)e.pc.kappa, cc = (double)e.pc.two_kappa_sqrtpi;
        const double hz = (double)e.rc * (double)e.rc / COUL_TAB;
        auto H = [&](double z) { const double r = std::sqrt(z); return std::erfc(kap * r) + cc * r * std::exp(-kap * kap * z); };
        auto dH = [&](double z) {
            const double r = std::sqrt(z);
            if (r == 0.0) return 0.0;
            return (cc * (1.0 - 2.0 * kap * kap * z) - 2.0 * kap / std::sqrt(M_PI)) * std::exp(-kap * kap * z) / (2.0 * r);
        };
        std::vector<float4> tab(COUL_TAB + 1);
        for (int i = 0; i < COUL_TAB; i++) {
            const double z0 = i * hz, z1 = (i + 1) * hz;
            const double h0 = H(z0), h1 = H(z1), d0 = dH(z0) * hz, d1 = dH(z1) * hz;
            tab[i] = make_float4((float)h0, (float)d0, (float)(3.0 * (h1 - h0) - 2.0 * d0 - d1), (float)(2.0 * (h0 - h1) + d0 + d1));
        }
        tab[COUL_TAB] = make_float4((float)H(COUL_TAB * hz), 0.f, 0.f, 0.f);   // r2 == cutoff^2 exactly
        e.coul_tab = dalloc_copy(tab);
    }
    e.dt_ns = (double)(p->dt_fs * 1e-6f);
    for (int d = 0; d < 3; d++) {
        e.Lq[d] = llround(b->box_nm[d] * 1e8);
        e.L[d] = (double)e.Lq[d] / 1e8;
        e.Lf[d] = (float)e.L[d];
        if (e.Lq[d] <= 0 || e.Lq[d] >= (1ll << 32)) throw LimaFail(LIMA_ERR_BAD_INPUT, "box edge must be in (0, 42.9) nm");
        if (e.L[d] <= (double)e.rc + e.skin) throw LimaFail(LIMA_ERR_BAD_INPUT, "box edge must exceed cutoff + skin");
    }
    // ---- global particle order ----
    e.h_first.assign(e.nC + 1, 0);
    for (int c = 0; c < e.nC; c++) {
        const int n = b->compound_n_particles[c];
        if (n < 0 || n > 32) throw LimaFail(LIMA_ERR_BAD_INPUT, "compound particle count out of [0, 32]");
        e.h_first[c + 1] = e.h_first[c] + n;
    }
    e.N = e.h_first[e.nC] + e.nT;
    if (e.N <= 0) throw LimaFail(LIMA_ERR_BAD_INPUT, "empty box");
    if (e.N >= (1 << 23)) throw LimaFail(LIMA_ERR_BAD_INPUT, "at most 2^23 particles");
    const int64_t N = e.N;
    std::vector<uint4> w(N);
    std::vector<int4> img(N);
    std::vector<float4> v(N), prm(N), zero4(N, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<int> comp(N);
    std::vector<float> qv(N);
    e.h_key.resize(e.nC);
    auto put = [&](int64_t i, const int64_t* u) {
        uint32_t ww[3];
        int im[3];
        for (int d = 0; d < 3; d++) {
            const int64_t q = fdiv(u[d], e.Lq[d]);
            ww[d] = (uint32_t)(u[d] - q * e.Lq[d]);
            im[d] = (int)q;
        }
        w[i] = make_uint4(ww[0], ww[1], ww[2], 0);
        img[i] = make_int4(im[0], im[1], im[2], 0);
    };
    for (int c = 0; c < e.nC; c++) {
        e.h_key[c] = b->compound_key_particle ? b->compound_key_particle[c] : 0;
        for (int s = 0; s < b->compound_n_particles[c]; s++) {
            const int64_t i = e.h_first[c] + s;
            const size_t ps = (size_t)c * 32 + s;
            int64_t u[3];
            for (int d = 0; d < 3; d++) u[d] = (int64_t)b->compound_origo[c * 3 + d] * 100000000ll + b->particle_coord[ps * 3 + d];
            put(i, u);
            const float* pv = b->particle_velocity ? &b->particle_velocity[ps * 3] : nullptr;
            v[i] = make_float4(pv ? pv[0] : 0.f, pv ? pv[1] : 0.f, pv ? pv[2] : 0.f, 0.f);
            const float q = e.elec ? b->particle_charge[ps] : 0.f;
            const float m = b->particle_mass[ps];
            if (!(m > 0.f)) throw LimaFail(LIMA_ERR_BAD_INPUT, "particle mass must be positive");
            if (b->particle_sigma[ps] < 0.f || b->particle_epsilon[ps] < 0.f) throw LimaFail(LIMA_ERR_BAD_INPUT, "negative LJ parameter");
            prm[i] = make_float4(q, 0.5f * b->particle_sigma[ps], sqrtf(b->particle_epsilon[ps]), m);
            comp[i] = c;
            qv[i] = q;
        }
    }
    for (int t = 0; t < e.nT; t++) {
        const int64_t i = e.h_first[e.nC] + t;
        int64_t u[3];
        for (int d = 0; d < 3; d++) u[d] = (int64_t)b->tinymol_block[t * 3 + d] * 100000000ll + b->tinymol_coord[t * 3 + d];
        put(i, u);
        const float* pv = b->tinymol_velocity ? &b->tinymol_velocity[t * 3] : nullptr;
        v[i] = make_float4(pv ? pv[0] : 0.f, pv ? pv[1] : 0.f, pv ? pv[2] : 0.f, 0.f);
        if (!(b->tinymol_mass[t] > 0.f)) throw LimaFail(LIMA_ERR_BAD_INPUT, "tinymol mass must be positive");
        prm[i] = make_float4(0.f, 0.5f * b->tinymol_sigma[t], sqrtf(b->tinymol_epsilon[t]), b->tinymol_mass[t]);
        comp[i] = e.nC + t;
        qv[i] = 0.f;
    }
    e.h_comp = comp;
    e.p.n = N;
    e.p.w = dalloc_copy(w);
    e.p.img = dalloc_copy(img);
    e.p.v = dalloc_copy(v);
    e.p.fprev = dalloc_copy(zero4);
    e.p.prm = dalloc_copy(prm);
    e.p.comp = dalloc_copy(comp);
    HIP_CHECK(hipMalloc(&e.w_build, N * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&e.img_build, N * sizeof(int4)));
    HIP_CHECK(hipMalloc(&e.dmax2, sizeof(float)));
    HIP_CHECK(hipMemset(e.dmax2, 0, sizeof(float)));
    {   // thermostat scale 1 (Engine.cu:130-131), no temperature measured yet
        const float t0[2] = {1.f, 0.f};
        HIP_CHECK(hipMalloc(&e.thermo, sizeof(t0)));
        HIP_CHECK(hipMemcpy(e.thermo, t0, sizeof(t0), hipMemcpyHostToDevice));
    }
    HIP_CHECK(hipMalloc(&e.dpart, ((N + 63) / 64) * sizeof(float2)));
    HIP_CHECK(hipMemset(e.dpart, 0, ((N + 63) / 64) * sizeof(float2)));
    HIP_CHECK(hipMalloc(&e.w_try, N * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&e.img_try, N * sizeof(int4)));
    e.f_acc = dalloc_copy(zero4);
    HIP_CHECK(hipMalloc(&e.sd, sizeof(SdState)));
    // ---- compounds, exclusions, intra work items ----
    std::vector<uint8_t> bm(std::max(e.nC, 1), 0);
    const uint32_t* lut = b->bonded_lut;
    auto lutrow = [&](int c, int k, int i) -> uint32_t { return lut ? lut[((size_t)c * 5 + k) * 32 + i] : (k == 2 ? (1u << i) : 0u); };
    for (int c = 0; c < e.nC; c++) {
        bm[c] = 1u << 2;
        for (int k = 0; k < 5; k++) {
            const int B = c + k - 2;
            if (k == 2 || B < 0 || B >= e.nC) continue;
            bool any = false;
            for (int i = 0; i < 32; i++) any |= lutrow(c, k, i) != 0;
            if (any) bm[c] |= (uint8_t)(1u << k);
        }
    }
    std::vector<int4> items;
    std::vector<uint32_t> ilut;
    for (int A = 0; A < e.nC; A++) {
        const int nA = b->compound_n_particles[A];
        for (int k = 0; k < 5; k++) {
            if (!((bm[A] >> k) & 1)) continue;
            const int B = A + k - 2;
            const int nB = b->compound_n_particles[B];
            bool any = false;
            uint32_t rows[32];
            for (int i = 0; i < 32; i++) {
                uint32_t r = lutrow(A, k, i);
                if (A == B) r |= 1u << i;
                rows[i] = r;
                if (i < nA)
                    for (int j = 0; j < nB; j++) any |= !((r >> j) & 1u);
            }
            if (!any) continue;
            // (A, B, key particle of A, key particle of B): the keys fix B's image for A != B
            items.push_back(make_int4(A, B, (int)(e.h_first[A] + e.h_key[A]), (int)(e.h_first[B] + e.h_key[B])));
            for (int i = 0; i < 32; i++) ilut.push_back(rows[i]);
        }
    }
    e.n_intra = (int)items.size();
    e.intra_items = dalloc_copy(items);
    e.intra_lut = dalloc_copy(ilut);
    {   // one group per compound with items (its items are consecutive)
        std::vector<int2> groups;
        for (int k = 0; k < e.n_intra; k++) {
            if (groups.empty() || items[groups.back().x].x != items[k].x) groups.push_back(make_int2(k, 0));
            groups.back().y++;
        }
        e.n_intra_groups = (int)groups.size();
        e.intra_groups = dalloc_copy(groups);
    }
    e.first = dalloc_copy(e.h_first);
    e.bmask = dalloc_copy(bm);
    // ---- bonded terms ----
    auto gid = [&](int32_t id) -> int {
        const int c = id / 32, s = id % 32;
        if (id < 0 || c >= e.nC || s >= b->compound_n_particles[c]) throw LimaFail(LIMA_ERR_BAD_INPUT, "bonded particle id out of range");
        return (int)(e.h_first[c] + s);
    };
    {
        std::vector<int2> bid(b->n_bonds); std::vector<float2> bp(b->n_bonds);
        for (int k = 0; k < b->n_bonds; k++) {
            bid[k] = make_int2(gid(b->bond_ids[2 * k]), gid(b->bond_ids[2 * k + 1]));
            bp[k] = make_float2(b->bond_params[2 * k], b->bond_params[2 * k + 1]);
        }
        std::vector<int4> aid(b->n_angles); std::vector<float4> ap(b->n_angles);
        for (int k = 0; k < b->n_angles; k++) {
            aid[k] = make_int4(gid(b->angle_ids[3 * k]), gid(b->angle_ids[3 * k + 1]), gid(b->angle_ids[3 * k + 2]), 0);
            ap[k] = make_float4(b->angle_params[4 * k], b->angle_params[4 * k + 1], b->angle_params[4 * k + 2], b->angle_params[4 * k + 3]);
        }
        std::vector<int4> did(b->n_dihedrals); std::vector<float4> dp(b->n_dihedrals);
        for (int k = 0; k < b->n_dihedrals; k++) {
            did[k] = make_int4(gid(b->dihedral_ids[4 * k]), gid(b->dihedral_ids[4 * k + 1]), gid(b->dihedral_ids[4 * k + 2]),
                               gid(b->dihedral_ids[4 * k + 3]));
            dp[k] = make_float4(b->dihedral_params[3 * k], b->dihedral_params[3 * k + 1], b->dihedral_params[3 * k + 2], 0.f);
        }
        std::vector<int4> iid(b->n_impropers); std::vector<float2> ip(b->n_impropers);
        for (int k = 0; k < b->n_impropers; k++) {
            iid[k] = make_int4(gid(b->improper_ids[4 * k]), gid(b->improper_ids[4 * k + 1]), gid(b->improper_ids[4 * k + 2]),
                               gid(b->improper_ids[4 * k + 3]));
            ip[k] = make_float2(b->improper_params[2 * k], b->improper_params[2 * k + 1]);
        }
        std::vector<int2> pid(b->n_pairs); std::vector<float2> pp(b->n_pairs);
        for (int k = 0; k < b->n_pairs; k++) {
            pid[k] = make_int2(gid(b->pair_ids[2 * k]), gid(b->pair_ids[2 * k + 1]));
            pp[k] = make_float2(b->pair_params[2 * k], b->pair_params[2 * k + 1]);
        }
        // incidence lists: every atom of a term gets (type, role, term); atoms in global order
        {
            const int64_t nterms = (int64_t)b->n_bonds + b->n_angles + b->n_dihedrals + b->n_impropers + b->n_pairs;
            if (nterms >= INC_MAX_TERMS) throw LimaFail(LIMA_ERR_BAD_INPUT, "too many bonded terms");
            if (N >= (1 << 24)) throw LimaFail(LIMA_ERR_BAD_INPUT, "bonded lanes address at most 2^24 particles");
            std::vector<int> cnt(N + 1, 0);
            auto count = [&](int atom) { cnt[atom]++; };
            for (auto& t : bid) { count(t.x); count(t.y); }
            for (auto& t : aid) { count(t.x); count(t.y); count(t.z); }
            for (auto& t : did) { count(t.x); count(t.y); count(t.z); count(t.w); }
            for (auto& t : iid) { count(t.x); count(t.y); count(t.z); count(t.w); }
            for (auto& t : pid) { count(t.x); count(t.y); }
            std::vector<int> off(N + 1, 0);
            for (int64_t i = 0; i < N; i++) {
                if (cnt[i] > 256) throw LimaFail(LIMA_ERR_UNSUPPORTED, "a particle belongs to more than 256 bonded terms");
                off[i + 1] = off[i] + cnt[i];
            }
            std::vector<uint32_t> inc(off[N]);
            std::vector<int> fill(off.begin(), off.end() - 1);
            auto add = [&](int atom, int type, int role, int term) { inc[fill[atom]++] = inc_code(type, role, term); };
            for (int k = 0; k < b->n_bonds; k++) { add(bid[k].x, T_BOND, 0, k); add(bid[k].y, T_BOND, 1, k); }
            for (int k = 0; k < b->n_angles; k++) { add(aid[k].x, T_ANGLE, 0, k); add(aid[k].y, T_ANGLE, 1, k); add(aid[k].z, T_ANGLE, 2, k); }
            for (int k = 0; k < b->n_dihedrals; k++) {
                add(did[k].x, T_DIH, 0, k); add(did[k].y, T_DIH, 1, k); add(did[k].z, T_DIH, 2, k); add(did[k].w, T_DIH, 3, k);
            }
            for (int k = 0; k < b->n_impropers; k++) {
                add(iid[k].x, T_IMP, 0, k); add(iid[k].y, T_IMP, 1, k); add(iid[k].z, T_IMP, 2, k); add(iid[k].w, T_IMP, 3, k);
            }
            for (int k = 0; k < b->n_pairs; k++) { add(pid[k].x, T_PAIR, 0, k); add(pid[k].y, T_PAIR, 1, k); }
            // lanes: the incidences of an atom with at most 64 of them in consecutive lanes of one
            // wave; after them, each atom with 65..256 incidences in a 256-lane workgroup of its own,
            // one run of at most 64 per wave starting at the wave's first lane
            std::vector<uint32_t> lcode;
            std::vector<int> latom;
            auto pad_to = [&](size_t m) {
                while (lcode.size() % m) { lcode.push_back(LANE_PAD); latom.push_back((int)((lcode.size() - 1) & 63) << 24); }
            };
            for (int64_t i = 0; i < N; i++) {
                const int run = cnt[i];
                if (run == 0 || run > 64) continue;
                if ((int)(lcode.size() & 63) + run > 64) pad_to(64);
                const int last = (int)((lcode.size() + run - 1) & 63);
                for (int r = 0; r < run; r++) { lcode.push_back(inc[off[i] + r]); latom.push_back((int)i | (last << 24)); }
            }
            pad_to(256);
            e.bd.multi_lane0 = (int)lcode.size();
            for (int64_t i = 0; i < N; i++) {
                if (cnt[i] <= 64) continue;
                for (int q = off[i]; q < off[i + 1]; q += 64) {
                    const int run = std::min(64, off[i + 1] - q);
                    for (int r = 0; r < run; r++) { lcode.push_back(inc[q + r]); latom.push_back((int)i | ((run - 1) << 24)); }
                    pad_to(64);
                }
                pad_to(256);
            }
            e.bd.n_lanes = (int)lcode.size();
            e.bd.lane_code = dalloc_copy(lcode);
            e.bd.lane_atom = dalloc_copy(latom);
        }
        e.bd.nb = b->n_bonds; e.bd.na = b->n_angles; e.bd.nd = b->n_dihedrals; e.bd.ni = b->n_impropers; e.bd.np = b->n_pairs;
        e.bd.bid = dalloc_copy(bid); e.bd.bp = dalloc_copy(bp);
        e.bd.aid = dalloc_copy(aid); e.bd.ap = dalloc_copy(ap);
        e.bd.did = dalloc_copy(did); e.bd.dp = dalloc_copy(dp);
        e.bd.iid = dalloc_copy(iid); e.bd.ip = dalloc_copy(ip);
        e.bd.pid = dalloc_copy(pid); e.bd.pp = dalloc_copy(pp);
    }
    HIP_CHECK(hipMalloc(&e.acc, ACC_N * sizeof(double)));
    HIP_CHECK(hipMemset(e.acc, 0, ACC_N * sizeof(double)));
    HIP_CHECK(hipMalloc(&e.acc_part, REDUCE_BLOCKS * ACC_N * sizeof(double)));
    HIP_CHECK(hipMalloc(&e.fmax, sizeof(float)));
    // status words, their pinned host copies and the snapshot buffers
    HIP_CHECK(hipMalloc(&e.dflags, DF_N * sizeof(int)));
    HIP_CHECK(hipMemset(e.dflags, 0, DF_N * sizeof(int)));
    HIP_CHECK(hipHostMalloc((void**)&e.hflags, 2 * DF_N * sizeof(int), hipHostMallocDefault));
    std::memset(e.hflags, 0, 2 * DF_N * sizeof(int));
    HIP_CHECK(hipMalloc(&e.v_snap, N * sizeof(float4)));
    HIP_CHECK(hipMalloc(&e.fprev_snap, N * sizeof(float4)));
    HIP_CHECK(hipMalloc(&e.thermo_snap, 2 * sizeof(float)));
    HIP_CHECK(hipMalloc(&e.lstats_snap, 2 * sizeof(int)));
    HIP_CHECK(hipMalloc(&e.acc_snap, ACC_N * sizeof(double)));
    HIP_CHECK(hipStreamCreate(&e.s_main));
    HIP_CHECK(hipStreamCreate(&e.s_pme));
    HIP_CHECK(hipEventCreateWithFlags(&e.ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&e.ev_join, hipEventDisableTiming));
    for (auto& ev : e.ev_rebuild) HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    for (auto& ev : e.tev) HIP_CHECK(hipEventCreate(&ev));
    nlist_alloc(e);
    pme_alloc(e, qv, w);
    // energy records: regions sized by capacities, fixed for the engine's lifetime
    e.ereg = energy_regions(e);
    e.epart_cap = e.ereg.total;
    HIP_CHECK(hipMalloc(&e.epart, e.epart_cap * ACC_N * sizeof(double)));
    HIP_CHECK(hipMemset(e.epart, 0, e.epart_cap * ACC_N * sizeof(double)));
    HIP_CHECK(hipMalloc(&e.lstats, 2 * sizeof(int)));
    HIP_CHECK(hipMemset(e.lstats, 0, 2 * sizeof(int)));
    if (p->log_capacity > 0) {   // data log ring
        e.log_cap = p->log_capacity;
        e.log_P = (int64_t)e.nC * LIMA_MAX_COMPOUND_PARTICLES + e.nT;
        const int64_t bytes = (int64_t)e.log_cap * e.log_P * 2 * (int64_t)sizeof(float4);
        if (bytes > (int64_t)64 << 30) throw LimaFail(LIMA_ERR_BAD_INPUT, "the data log would exceed 64 GiB; lower log_capacity");
        std::vector<int> lidx(N), lkey(N, -1);
        for (int c = 0; c < e.nC; c++)
            for (int64_t i = e.h_first[c]; i < e.h_first[c + 1]; i++) {
                lidx[i] = c * LIMA_MAX_COMPOUND_PARTICLES + (int)(i - e.h_first[c]);
                lkey[i] = (int)(e.h_first[c] + e.h_key[c]);
            }
        for (int t = 0; t < e.nT; t++) lidx[e.h_first[e.nC] + t] = e.nC * LIMA_MAX_COMPOUND_PARTICLES + t;
        e.log_index = dalloc_copy(lidx);
        e.log_key = dalloc_copy(lkey);
        HIP_CHECK(hipMalloc(&e.log_pe, bytes / 2));
        HIP_CHECK(hipMalloc(&e.log_fv, bytes / 2));
        HIP_CHECK(hipMemset(e.log_pe, 0, bytes / 2));
        HIP_CHECK(hipMemset(e.log_fv, 0, bytes / 2));
        HIP_CHECK(hipMalloc(&e.log_acc, (size_t)e.log_cap * ACC_N * sizeof(double)));
        e.log_step.assign(e.log_cap, -1);
    }
    nlist_build(e, e.p.w, /*checked=*/true);
    HIP_CHECK(hipStreamSynchronize(e.s_main));
    e.energy.step = -1;
}

void destroy(Engine& e) {
    if (e.s_main) (void)hipStreamSynchronize(e.s_main);
    nlist_free(e);
    pme_free(e);
    void* ptrs[] = {e.p.w, e.p.img, e.p.v, e.p.fprev, e.p.prm, e.p.comp, e.w_build, e.img_build, e.dmax2, e.lstats, e.thermo, e.dpart, e.epart, e.w_try, e.img_try, e.f_acc, e.sd, e.first,
                    e.bmask, e.intra_items, e.intra_lut, e.intra_groups, e.bd.bid, e.bd.bp, e.bd.aid, e.bd.ap, e.bd.did, e.bd.dp,
                    e.bd.iid, e.bd.ip, e.bd.pid, e.bd.pp, e.bd.lane_code, e.bd.lane_atom, e.coul_tab, e.acc, e.acc_part, e.fmax,
                    e.log_pe, e.log_fv, e.log_acc, e.log_index, e.log_key, e.dflags, e.v_snap, e.fprev_snap, e.thermo_snap,
                    e.lstats_snap, e.acc_snap};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    if (e.hflags) (void)hipHostFree(e.hflags);
    for (auto& ev : e.ev_rebuild) if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : e.tev) if (ev) (void)hipEventDestroy(ev);
    if (e.step_graph) (void)hipGraphExecDestroy(e.step_graph);
    e.step_graph = nullptr;
    if (e.ev_fork) (void)hipEventDestroy(e.ev_fork);
    if (e.ev_join) (void)hipEventDestroy(e.ev_join);
    if (e.s_main) (void)hipStreamDestroy(e.s_main);
    if (e.s_pme) (void)hipStreamDestroy(e.s_pme);
}

// Forces (and potE) of the last evaluation in global order into out, max |F| into e.fmax.
void forces_global(Engine& e, float4* out) {
    hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, e.s_main, e.fmax, 1);
    hipLaunchKernelGGL(k_forces_global, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, e.N, e.nl.slot_of, slot_forces(e), out,
                       e.fmax);
}

}  // namespace

extern "C" {

const char* lima_last_error(void) { return g_err.c_str(); }

lima_status lima_create(const lima_box_desc* box, const lima_simparams* params, lima_engine** out) {
    if (!out) { set_err("null output handle"); return LIMA_ERR_BAD_INPUT; }
    *out = nullptr;
    auto* h = new lima_engine();
    const lima_status st = guarded([&] { create(h->e, box, params); });
    if (st != LIMA_OK) {
        destroy(h->e);
        delete h;
        return st;
    }
    *out = h;
    return LIMA_OK;
}

lima_status lima_destroy(lima_engine* h) {
    if (!h) return LIMA_OK;
    destroy(h->e);
    delete h;
    return LIMA_OK;
}

lima_status lima_step(lima_engine* h, int64_t nsteps) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    if (h->e.critical) { set_err("engine is in a critical error state"); return LIMA_ERR_NONFINITE; }
    if (nsteps < 0) { set_err("negative step count"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] { issue_steps(h->e, h->e.step + nsteps); });
}

lima_status lima_synchronize(lima_engine* h) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] { settle(h->e); });
}

// One evaluation with energies at the current positions into e.energy; a PME tile bin that
// overflows (the only capacity an evaluation can exceed after a checked rebuild) is grown and the
// evaluation repeated.
void compute_forces_now(Engine& e, bool with_kinetic) {
    for (int attempt = 0;; attempt++) {
        nlist_build(e, e.p.w, /*checked=*/true);
        evaluate(e, e.p.w, e.p.img, true);
        if (with_kinetic)
            hipLaunchKernelGGL(k_kinetic, dim3(nblk(e.N, 256)), dim3(256), 0, e.s_main, e.N, e.p.v, e.p.prm,
                               e.epart + (int64_t)e.ereg.kin * ACC_N);
        launch_reduce_energy(e, with_kinetic, e.s_main);
        HIP_CHECK(hipStreamSynchronize(e.s_main));
        HIP_CHECK(hipGetLastError());
        int f[DF_N];
        HIP_CHECK(hipMemcpy(f, e.dflags, sizeof(f), hipMemcpyDeviceToHost));
        if (!f[DF_PMEBIN]) break;
        if (attempt >= 6) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "PME tile bin capacity exceeded");
        pme_grow_bins(e);
        reset_flags(e, f[DF_SNAPSHOTS], e.s_main);
        pme_clear_bins(e, e.s_main);
    }
    e.packed = false;
}

lima_status lima_compute_forces(lima_engine* h, lima_energy* out) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        compute_forces_now(e, true);
        fill_energy(e, true, e.step);
        e.energy_step = e.step;
        e.forces_valid = true;
        if (out) *out = e.energy;
    });
}

lima_status lima_get_energy(lima_engine* h, lima_energy* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        settle(h->e);
        *out = h->e.energy;
    });
}

lima_status lima_minimize(lima_engine* h, double tol_kj, int64_t max_iter, double h0, lima_em_result* res) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = e.s_main;
        // initial evaluation at the current positions
        compute_forces_now(e, false);
        forces_global(e, e.f_acc);
        SdState sd{};
        {
            double acc[ACC_N];
            float fm;
            HIP_CHECK(hipMemcpyAsync(acc, e.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(&fm, e.fmax, sizeof(float), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            sd.e_acc = acc[ACC_LJ] + acc[ACC_COUL] + acc[ACC_BOND] + acc[ACC_PME] + e_const(e);
            sd.fmax_acc = fm;
            sd.h = (float)h0;
        }
        HIP_CHECK(hipMemcpyAsync(e.sd, &sd, sizeof(sd), hipMemcpyHostToDevice, s));
        const double tol = tol_kj * 1000.0;
        int64_t it = 0;
        bool conv = sd.fmax_acc < tol;
        while (!conv && it < max_iter) {
            if (sd.disp_since_build + sd.h > 0.5 * e.skin) {
                nlist_build(e, e.p.w, /*checked=*/true);
                sd.disp_since_build = 0;
                HIP_CHECK(hipMemcpyAsync(&e.sd->disp_since_build, &sd.disp_since_build, sizeof(double), hipMemcpyHostToDevice, s));
            }
            hipLaunchKernelGGL(k_sd_trial, dim3(nblk(e.N, 256)), dim3(256), 0, s, e.N, (long long)e.Lq[0], (long long)e.Lq[1],
                               (long long)e.Lq[2], e.sd, e.p.w, e.p.img, e.f_acc, e.w_try, e.img_try);
            evaluate(e, e.w_try, e.img_try, true);
            launch_reduce_energy(e, false, s);
            // trial forces in global order -> p.fprev used as scratch
            forces_global(e, e.p.fprev);
            hipLaunchKernelGGL(k_sd_accept, dim3(1), dim3(1), 0, s, e.sd, e.acc, e_const(e), e.fmax, (const int*)e.dflags);
            hipLaunchKernelGGL(k_sd_commit, dim3(nblk(e.N, 256)), dim3(256), 0, s, e.N, e.sd, e.w_try, e.img_try, e.p.fprev,
                               e.p.w, e.p.img, e.f_acc);
            int f[DF_N];
            HIP_CHECK(hipMemcpyAsync(&sd, e.sd, sizeof(sd), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(f, e.dflags, sizeof(f), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (f[DF_PMEBIN]) {   // the trial was refused without a change of h: grow the bins and repeat it
                if (e.pme.bincap > (1 << 20)) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "PME tile bin capacity exceeded");
                pme_grow_bins(e);
                reset_flags(e, f[DF_SNAPSHOTS], s);
                pme_clear_bins(e, s);
                continue;
            }
            it++;
            conv = sd.fmax_acc < tol;
            if (!std::isfinite(sd.e_acc)) throw LimaFail(LIMA_ERR_NONFINITE, "non-finite energy in minimisation");
            if (sd.h < 1e-12f) break;
        }
        // leave forces / energies of the accepted state in the engine, with a list built at the
        // final positions (MD that follows starts with its full skin)
        compute_forces_now(e, false);
        HIP_CHECK(hipMemsetAsync(e.p.fprev, 0, e.N * sizeof(float4), s));
        HIP_CHECK(hipStreamSynchronize(s));
        fill_energy(e, false, e.step);
        e.energy_step = e.step;
        e.forces_valid = true;
        if (res) {
            res->iterations = it;
            res->accepted = sd.n_accepted;
            res->final_energy = sd.e_acc;
            res->final_fmax = sd.fmax_acc / 1000.0;
            res->converged = conv ? 1 : 0;
            res->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}

lima_status lima_get_state(lima_engine* h, lima_state_out* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        const int64_t N = e.N;
        std::vector<uint4> w(N);
        std::vector<int4> img(N);
        std::vector<float4> v(N), f(N);
        HIP_CHECK(hipMemcpy(w.data(), e.p.w, N * sizeof(uint4), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(img.data(), e.p.img, N * sizeof(int4), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(v.data(), e.p.v, N * sizeof(float4), hipMemcpyDeviceToHost));
        if (e.forces_valid) {
            forces_global(e, e.f_acc);
            HIP_CHECK(hipStreamSynchronize(e.s_main));
            HIP_CHECK(hipMemcpy(f.data(), e.f_acc, N * sizeof(float4), hipMemcpyDeviceToHost));
        } else {
            std::fill(f.begin(), f.end(), make_float4(0.f, 0.f, 0.f, 0.f));
        }
        auto unw = [&](int64_t i, int d) -> int64_t {
            const uint32_t ww = d == 0 ? w[i].x : (d == 1 ? w[i].y : w[i].z);
            const int im = d == 0 ? img[i].x : (d == 1 ? img[i].y : img[i].z);
            return (int64_t)ww + (int64_t)im * e.Lq[d];
        };
        auto wr = [&](int64_t i, int d) -> int64_t { return d == 0 ? w[i].x : (d == 1 ? w[i].y : w[i].z); };
        for (int c = 0; c < e.nC; c++) {
            const int64_t k = e.h_first[c] + e.h_key[c];
            int64_t base[3];
            for (int d = 0; d < 3; d++) {
                const int64_t wk = wr(k, d);
                const int32_t org = (int32_t)(wk / 100000000ll);
                if (out->compound_origo) out->compound_origo[c * 3 + d] = org;
                base[d] = unw(k, d) - wk + (int64_t)org * 100000000ll;
            }
            for (int s = 0; s < (int)(e.h_first[c + 1] - e.h_first[c]); s++) {
                const int64_t i = e.h_first[c] + s;
                const size_t ps = (size_t)c * 32 + s;
                const float vv[3] = {v[i].x, v[i].y, v[i].z}, ff[3] = {f[i].x, f[i].y, f[i].z};
                for (int d = 0; d < 3; d++) {
                    if (out->particle_coord) out->particle_coord[ps * 3 + d] = (int32_t)(unw(i, d) - base[d]);
                    if (out->particle_velocity) out->particle_velocity[ps * 3 + d] = vv[d];
                    if (out->particle_force) out->particle_force[ps * 3 + d] = ff[d];
                }
                if (out->particle_pot_energy) out->particle_pot_energy[ps] = f[i].w;
            }
        }
        for (int t = 0; t < e.nT; t++) {
            const int64_t i = e.h_first[e.nC] + t;
            const float vv[3] = {v[i].x, v[i].y, v[i].z}, ff[3] = {f[i].x, f[i].y, f[i].z};
            for (int d = 0; d < 3; d++) {
                const int64_t ww = wr(i, d);
                const int32_t blk = (int32_t)(ww / 100000000ll);
                if (out->tinymol_block) out->tinymol_block[t * 3 + d] = blk;
                if (out->tinymol_coord) out->tinymol_coord[t * 3 + d] = (int32_t)(ww - (int64_t)blk * 100000000ll);
                if (out->tinymol_velocity) out->tinymol_velocity[t * 3 + d] = vv[d];
                if (out->tinymol_force) out->tinymol_force[t * 3 + d] = ff[d];
            }
            if (out->tinymol_pot_energy) out->tinymol_pot_energy[t] = f[i].w;
        }
    });
}

lima_status lima_get_log_info(lima_engine* h, lima_log_info* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        out->frames_logged = e.log_frames;
        out->first_held = e.log_frames - std::min<int64_t>(e.log_frames, e.log_cap);
        out->particles_per_frame = (int64_t)e.nC * LIMA_MAX_COMPOUND_PARTICLES + e.nT;
        out->interval = e.prm.data_logging_interval;
        out->capacity = e.log_cap;
    });
}

lima_status lima_read_log(lima_engine* h, int64_t first, int64_t n, lima_log_out* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        const int64_t held0 = e.log_frames - std::min<int64_t>(e.log_frames, e.log_cap);
        if (n < 0 || first < held0 || first + n > e.log_frames)
            throw LimaFail(LIMA_ERR_BAD_INPUT, "frames " + std::to_string(first) + " + " + std::to_string(n) + " are not held (held: " +
                                                   std::to_string(held0) + " .. " + std::to_string(e.log_frames - 1) + ")");
        HIP_CHECK(hipStreamSynchronize(e.s_main));
        const int64_t P = e.log_P;
        std::vector<float4> pe(P), fv(P);
        for (int64_t k = 0; k < n; k++) {
            const int64_t slot = (first + k) % e.log_cap;
            if (out->step) out->step[k] = e.log_step[slot];
            if (out->energy) {
                double acc[ACC_N];
                HIP_CHECK(hipMemcpy(acc, e.log_acc + slot * ACC_N, sizeof(acc), hipMemcpyDeviceToHost));
                lima_energy E{};
                energy_from_acc(e, acc, true, E);
                out->energy[k] = E;
            }
            if (out->position || out->pot_energy)
                HIP_CHECK(hipMemcpy(pe.data(), e.log_pe + slot * P, P * sizeof(float4), hipMemcpyDeviceToHost));
            if (out->force || out->speed)
                HIP_CHECK(hipMemcpy(fv.data(), e.log_fv + slot * P, P * sizeof(float4), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < P; i++) {
                const size_t o = (size_t)(k * P + i);
                if (out->position) { out->position[3 * o] = pe[i].x; out->position[3 * o + 1] = pe[i].y; out->position[3 * o + 2] = pe[i].z; }
                if (out->pot_energy) out->pot_energy[o] = pe[i].w;
                if (out->force) { out->force[3 * o] = fv[i].x; out->force[3 * o + 1] = fv[i].y; out->force[3 * o + 2] = fv[i].z; }
                if (out->speed) out->speed[o] = fv[i].w;
            }
        }
    });
}

lima_status lima_get_status(lima_engine* h, lima_runstatus* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    Engine& e = h->e;
    std::memset(out, 0, sizeof(*out));
    const lima_status status = e.critical ? LIMA_OK : guarded([&] { settle(e); });
    out->current_step = e.step;
    out->current_temperature = NAN;   // RunStatus default until the first measurement (Engine.cuh:67)
    if (e.temp_measured) {
        float t[2];
        if (hipMemcpy(t, e.thermo, sizeof(t), hipMemcpyDeviceToHost) == hipSuccess) out->current_temperature = t[1];
    }
    out->critical_error_occured = e.critical ? 1 : 0;
    out->simulation_finished = (e.prm.n_steps > 0 && e.step >= e.prm.n_steps) ? 1 : 0;
    out->n_list_rebuilds = e.n_rebuilds;
    int ncl = 0;
    if (e.nl.ncl_dev && hipMemcpy(&ncl, e.nl.ncl_dev, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess && ncl > 0) {
        e.nl.ncl = ncl;   // padded entry slots of the current list (summed here, not per rebuild)
        std::vector<int> cnt(ncl), far(ncl);
        if (hipMemcpy(cnt.data(), e.nl.ecount, cnt.size() * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(far.data(), e.nl.efar, far.size() * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess)
            for (int c = 0; c < ncl; c++) {
                out->list_entries += cnt[c] + far[c];
                out->list_max_entries = std::max(out->list_max_entries, cnt[c] + far[c]);
            }
    }
    int st[2] = {0, 0};
    if (e.lstats && hipMemcpy(st, e.lstats, sizeof(st), hipMemcpyDeviceToHost) == hipSuccess) {
        out->list_violations = st[0];
        float d;
        std::memcpy(&d, &st[1], sizeof(float));
        out->max_displacement_nm = d;
    }
    out->greatest_force = NAN;
    out->capacity_rollbacks = e.rollbacks;
    out->list_interval = e.list_interval;
    out->list_stale_rollbacks = (int32_t)e.stale_rollbacks;
    out->list_capacity = e.nl.cap;
    return status;
}

// The current list read back in cluster order: per cluster its near, then its far entries, mask-0
// padding dropped; entries as global particle | shift code << 24.
struct HostList {
    std::vector<int> catoms;
    std::vector<int64_t> off;
    std::vector<uint32_t> ent;
    std::vector<uint8_t> mask;
};
static HostList read_list(Engine& e) {
    HostList L;
    const int ncl = nlist_count(e);
    const int64_t cap = e.nl.cap;
    std::vector<int> cnt(ncl), far(ncl);
    std::vector<uint2> ent((size_t)ncl * cap);
    L.catoms.resize((size_t)ncl * CLUSTER);
    HIP_CHECK(hipMemcpy(L.catoms.data(), e.nl.catoms, L.catoms.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (ncl) {
        HIP_CHECK(hipMemcpy(cnt.data(), e.nl.ecount, ncl * sizeof(int), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(far.data(), e.nl.efar, ncl * sizeof(int), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(ent.data(), e.nl.ent, ent.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    }
    L.off.assign(ncl + 1, 0);
    for (int I = 0; I < ncl; I++) {
        L.off[I] = (int64_t)L.ent.size();
        auto take = [&](int64_t k) {
            const uint32_t bits = ent[k].x, m = bits >> 24;
            if (!m) return;
            L.ent.push_back((uint32_t)L.catoms[bits & 0xffffffu] | ((ent[k].y & 0xffu) << 24));
            L.mask.push_back((uint8_t)m);
        };
        for (int64_t k = I * cap; k < I * cap + cnt[I]; k++) take(k);
        for (int64_t k = I * cap + cap - far[I]; k < (I + 1) * cap; k++) take(k);
    }
    L.off[ncl] = (int64_t)L.ent.size();
    return L;
}

lima_status lima_get_neighbor_counts(lima_engine* h, int64_t* ncl, int64_t* nent) {
    if (!h || !ncl || !nent) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        nlist_build(e, e.p.w, /*checked=*/true);
        e.packed = false;
        HIP_CHECK(hipStreamSynchronize(e.s_main));
        const HostList L = read_list(e);
        *ncl = (int64_t)L.off.size() - 1;
        *nent = (int64_t)L.ent.size();
    });
}

lima_status lima_get_neighbor_list(lima_engine* h, int32_t* cluster_atoms, int64_t* offsets, uint32_t* entries, uint8_t* masks) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        Engine& e = h->e;
        settle(e);
        const HostList L = read_list(e);
        if (cluster_atoms) std::memcpy(cluster_atoms, L.catoms.data(), L.catoms.size() * sizeof(int));
        if (offsets) std::memcpy(offsets, L.off.data(), L.off.size() * sizeof(int64_t));
        if (entries) std::memcpy(entries, L.ent.data(), L.ent.size() * sizeof(uint32_t));
        if (masks) std::memcpy(masks, L.mask.data(), L.mask.size());
    });
}

lima_status lima_get_timings(lima_engine* h, lima_timings* out) {
    if (!h || !out) { set_err("null argument"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        settle(h->e);
        timing_flush(h->e);
        *out = h->e.tacc;
        out->host_issue_ms = h->e.host_issue_ms;
        out->host_steps = h->e.host_steps;
    });
}

lima_status lima_set_timing(lima_engine* h, int32_t enabled) {
    if (!h) { set_err("null engine"); return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        settle(h->e);
        timing_flush(h->e);
        h->e.timing = enabled != 0;
        h->e.tacc = lima_timings{};
    });
}

int64_t lima_get_pair_count(lima_engine* h) {
    if (!h) return -1;
    int64_t n = -1;
    const lima_status st = guarded([&] {
        settle(h->e);
        n = h->e.pairs_last;
    });
    return st == LIMA_OK ? n : -1;
}

}  // extern "C"
