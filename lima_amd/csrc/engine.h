// engine.h -- device state of one lima_engine and the launchers each .hip file provides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lima_hip.h"
#include "lima_dev.h"
#include "lima_terms.h"

namespace lima {

// Arrays indexed by global particle (compound particles in compound order, then tinymols).
struct Particles {
    int64_t n = 0;
    uint4* w = nullptr;        // wrapped Coord [0, Lq)
    int4* img = nullptr;       // periodic image counters: unwrapped = w + img * Lq
    float4* v = nullptr;       // velocity nm/ns
    float4* fprev = nullptr;   // force of the previous step
    float4* prm = nullptr;     // (charge kC/mol, sigma/2 nm, sqrt(eps J/mol), mass kg/mol)
    int* comp = nullptr;       // compound id (tinymols: n_compounds + t)
};

// Cluster-pair neighbour list (definition in oracle/lima_oracle.cpp build_clusters).
struct NList {
    int ncx = 1, ncy = 1, ncol = 1;
    int ncl = 0;               // clusters of the current list (host copy, refreshed by nlist_count)
    int ncl_max = 0;           // capacity: every grid, slot array and energy region is sized by it, and
                               // kernels read the current count from ncl_dev, so a rebuild changes no
                               // kernel argument
    const int* ncl_dev = nullptr;   // = col_count + ncol after a rebuild (device)
    int64_t nent_cap = 0;      // entry slots allocated (ncl_max * cap)
    int cap = 2048;            // entry capacity per i-cluster (grows on overflow)
    int* col_of = nullptr;     // [N]
    int* col_count = nullptr;  // [ncol+1]
    int* col_start = nullptr;  // [ncol+1]
    int* col_fill = nullptr;   // [ncol]
    unsigned long long* keys = nullptr;  // [N] (w_z << 32 | particle)
    int* col_ncl = nullptr;    // [ncol+1] clusters per column -> exclusive scan
    int* catoms = nullptr;     // [ncl_max*8] particle or -1
    int* slot_of = nullptr;    // [N]
    int* ccol = nullptr;       // [ncl_max] column of cluster
    uint4* cref = nullptr;     // [ncl_max] cluster reference point (integer, centre of bbox)
    uint4* cmin = nullptr;     // [ncl_max]
    uint4* cmax = nullptr;     // [ncl_max]
    float2* ljs = nullptr;     // [slots] (sigma/2, sqrt(24 eps))
    int* comps = nullptr;      // [slots] compound of slot (-1 padding)
    uint4* cw = nullptr;       // [slots] wrapped Coord of the slot's particle (w = 1 for padding)
    long long* xb[2] = {nullptr, nullptr};   // [n+1] first Coord of each x / y column: ceil(c Lq / n)
    double rinv[2] = {0.0, 0.0};             // n / Lq per axis
    float4* xq = nullptr;      // [slots] position relative to cluster ref [nm] + charge
    // Each i-cluster owns the slice [I cap, (I+1) cap) of the entry arrays.  Entries whose j image lay
    // within the cutoff of the cluster's box at the build fill it from the front (ecount[I] of them,
    // padded to whole passes of 8), the farther ones (up to cutoff + skin) from the back (efar[I]):
    // the pair kernel's passes over the far segment mostly find every pair outside the cutoff and
    // skip their arithmetic.
    int* ecount = nullptr;     // [ncl_max+1] near entries (padded)
    int* efar = nullptr;       // [ncl_max+1] far entries (padded), at the slice's end
    // [nent_cap] x: j slot | mask << 24;  y: image shift code sx + 3 sy + 9 sz (each 0..2 for -1, 0, +1
    // box) in bits 0-7, and per axis d a -1 box bit (8 + 2d) and a +1 box bit (9 + 2d).  The pair
    // kernel forms the shift between the two clusters' reference points from these bits and cref.
    uint2* ent = nullptr;
};

struct Pme {
    int K[3] = {0, 0, 0};
    int H = 0;                             // K[0]/2 + 1
    long long* qgrid = nullptr;            // [Kz][Ky][Kx] fixed point charge * weight
    float2* cgrid = nullptr;               // [Kz][Ky][H]
    float* phi = nullptr;                  // [Kz][Ky][Kx]
    float* green = nullptr;                // [Kz][Ky][H]
    float2* tw[3] = {nullptr, nullptr, nullptr};   // twiddles exp(-2 pi i k / K_d)
    int radix[3][8];
    int nrad[3] = {0, 0, 0};
    double eself_total = 0;                // -kappa/sqrt(pi_L) kC sum q^2
    int64_t ncharged = 0;
    // tiled spreading: charged particles binned every step by the tile of their base node;
    // each tile accumulates an E^3 block (E = T + 3) in LDS
    bool tiled = false;
    int T[3] = {8, 8, 8}, margin = 1, E[3] = {13, 13, 13};   // tile edge divides K; E = T + 3 + 2 margin
    int nt[3] = {0, 0, 0};
    int ntiles = 0;
    int bincap = 512;                      // charged particles per tile bin
    int* bin_count = nullptr;              // [ntiles+1]
    int* bin_slot = nullptr;               // [ntiles][bincap] slot of each binned particle
    uint4* bin_pos = nullptr;              // [ntiles][bincap] its wrapped Coord and charge bits
    float* tblock = nullptr;               // [ntiles][E^3] per-tile blocks (charge density, float)
    int4* cover[3] = {nullptr, nullptr, nullptr};   // per grid coordinate: (tile a, local a, tile b, local b)
    int* flags = nullptr;                  // the engine's DF_PMEBIN word: a bin overflowed / a particle left its block
};

struct Bonded {
    int nb = 0, na = 0, nd = 0, ni = 0, np = 0;
    int2* bid = nullptr; float2* bp = nullptr;
    int4* aid = nullptr; float4* ap = nullptr;
    int4* did = nullptr; float4* dp = nullptr;
    int4* iid = nullptr; float2* ip = nullptr;
    int2* pid = nullptr; float2* pp = nullptr;
    // incidence lanes: one lane per (atom, term) pair, the lanes of an atom contiguous and inside
    // one wave; lane_atom = atom | (last lane of the atom's run within the wave) << 24
    int n_lanes = 0;             // multiple of 64
    uint32_t* lane_code = nullptr;   // inc_code, or LANE_PAD
    int* lane_atom = nullptr;
    int multi_lane0 = 0;     // first lane of the workgroups that each hold one atom with 65..256 incidences
};
constexpr uint32_t LANE_PAD = 0xffffffffu;

// Device accumulators (FP64 energy reductions and scalars).
constexpr int LAUNCH_EAGER = LIMA_LAUNCH_EAGER, LAUNCH_ONE_STREAM = LIMA_LAUNCH_ONE_STREAM;

// ACC_VXX .. ACC_VYZ: the virial W = -dU/d(strain), components xx yy zz xy xz yz [J/mol]
enum Acc { ACC_LJ = 0, ACC_COUL, ACC_BOND, ACC_PME, ACC_KIN, ACC_PAIRS, ACC_VXX, ACC_VYY, ACC_VZZ, ACC_VXY, ACC_VXZ, ACC_VYZ,
           ACC_N };
constexpr int REDUCE_BLOCKS = 64;   // first stage of the energy reduction

// FP64 energy reduction without same-address atomics: every wave of an energy-producing kernel
// writes one record of ACC_N doubles into its kernel's region of `epart`; k_reduce_energy sums
// the records in a fixed order (deterministic totals).
struct ERegions {
    int pair = 0, intra = 0, bond = 0, pme = 0, kin = 0, total = 0;
};

// One energy record: a kernel's sums, zero in the fields it does not produce (vir: six virial
// components or nullptr).
__device__ __forceinline__ void write_record(double* r, double lj, double coul, double bond, double kin, double pairs,
                                             const double* vir) {
    r[ACC_LJ] = lj; r[ACC_COUL] = coul; r[ACC_BOND] = bond; r[ACC_PME] = 0.0; r[ACC_KIN] = kin; r[ACC_PAIRS] = pairs;
#pragma unroll
    for (int k = 0; k < 6; k++) r[ACC_VXX + k] = vir ? vir[k] : 0.0;
}

// Sticky status words on the device (Engine::dflags).  Kernels only set them; the host clears
// them after it has dealt with the condition.  Every list rebuild whose error words are still
// clear snapshots the MD state (positions, velocities, previous forces, thermostat scale) and
// counts the snapshot in DF_SNAPSHOTS, so a condition found later is undone by restoring the
// last snapshot, growing the capacity that overflowed and issuing the steps again.
enum DFlag {
    DF_COLUMN = 0,     // an x-y column of the cell grid holds more particles than the sort can take
    DF_PAIRLIST,       // an i-cluster's pair-list slice overflowed
    DF_PMEBIN,         // a PME tile bin overflowed
    DF_NONFINITE,      // the integrator met a NaN / inf force
    DF_STALE,          // a list interval ended with two displacements since the build summing to more
                       // than the skin: a pair could have come inside the cutoff unlisted
    DF_SNAPSHOTS,      // snapshots taken (ordinal of the last one)
    DF_LIST_NEED,      // with DF_PAIRLIST: the largest entry count an i-cluster needed
    DF_N = 8
};
constexpr int DF_ERRORS = 5;   // words 0 .. DF_ERRORS-1 are error conditions

// Host side of a snapshot: the engine counters that belong with the device state it saved.
struct SnapRecord {
    int64_t ordinal = 0;       // value DF_SNAPSHOTS takes when this snapshot is taken
    int64_t step = 0;
    int64_t log_frames = 0;
    int n_rebuilds = 0;        // rebuilds before this one
    bool temp_measured = false;
    int64_t energy_step = -1;
    int list_interval = 20;    // steps of the list this rebuild starts
    int builds_clean = 0;
};

struct SdState {            // steepest descent state, device resident
    double e_acc, e_try;
    float fmax_acc, fmax_try;
    float h;
    int accepted;
    long long n_accepted;
    double disp_since_build;
};

struct Engine {
    lima_simparams prm{};
    int nC = 0, nT = 0;
    int64_t N = 0;
    int64_t Lq[3]{};
    double L[3]{};
    float Lf[3]{};
    float rc = 1.2f, rc2 = 1.44f, skin = 0.1f;
    PairConst pc{};
    float4* coul_tab = nullptr;   // [COUL_TAB + 1] cubic Hermite coefficients of the Coulomb force factor in r^2
    double dt_ns = 2e-6;
    bool elec = true;

    Particles p;
    // SD trial state
    uint4* w_build = nullptr;  // positions at the last list rebuild (displacement check)
    int4* img_build = nullptr;
    float* dmax2 = nullptr;    // [1] max squared displacement since the last rebuild [nm^2]
    int* lstats = nullptr;     // [2] device list statistics: rebuilds after which 2 max|dx| > skin, last max|dx| (float bits)
    float2* dpart = nullptr;   // [N/64] per wave: running maxima of the largest and second-largest squared
                               // displacement since the list build among the wave's particles
    int list_interval = 20;        // steps per list in use: nlist_interval, lowered after a stale interval
    int builds_clean = 0;          // rebuilds issued since the last stale interval
    int grow_after = 20;           // clean rebuilds before lists are tried one step longer; doubles
                                   // (up to 640) with every stale interval, so a length that fails
                                   // now and then is retried ever more rarely
    int64_t stale_rollbacks = 0;   // intervals undone because the list could have missed a pair
    uint4* w_try = nullptr;
    int4* img_try = nullptr;
    float4* f_acc = nullptr;   // accepted forces, global order
    SdState* sd = nullptr;
    // snapshot of the MD state at the last clean list rebuild (positions: w_build / img_build)
    float4* v_snap = nullptr;
    float4* fprev_snap = nullptr;
    float* thermo_snap = nullptr;  // [2]
    int* lstats_snap = nullptr;    // [2]
    double* acc_snap = nullptr;    // [ACC_N]
    bool stats_skip = false;       // the next rebuild does not count list statistics (after a rollback)
    int* dflags = nullptr;         // [DF_N] sticky status words (device)
    int* hflags = nullptr;         // [2][DF_N] pinned host copies, written after each deferred rebuild
    hipEvent_t ev_rebuild[2] = {nullptr, nullptr};   // recorded after each deferred rebuild's flag copy
    bool rb_pending[2] = {false, false};
    int rb_parity = 0;
    int64_t snap_issued = 0;       // snapshots the host has asked for (device count if none was skipped)
    std::vector<SnapRecord> snaps; // records of snapshots not yet known to be superseded
    int rollbacks = 0;             // rollbacks since create (lima_runstatus.capacity_rollbacks)

    // compounds
    int64_t* first = nullptr;  // [nC+1]
    uint8_t* bmask = nullptr;  // [nC]
    std::vector<int64_t> h_first;
    std::vector<int> h_key;
    std::vector<int> h_comp;   // global -> compound
    // intra pairs work list
    int n_intra = 0;
    int4* intra_items = nullptr;    // (A, B, first A, first B)
    uint32_t* intra_lut = nullptr;  // [n_intra*32]
    int n_intra_groups = 0;         // compounds A with at least one item (their items are consecutive)
    int2* intra_groups = nullptr;   // (first item, item count) per group

    Bonded bd;
    NList nl;
    Pme pme;

    // Per-kernel force buffers, one (fx, fy, fz, potE share) record per slot, each written with plain
    // stores by the one wave that owns the slot; the integrator and the read-back sum them in a fixed
    // order (pair + intra + bonded + PME), so forces are deterministic.  Slots a kernel never writes
    // (no bonded terms, no intra pairs, no charge) are zeroed at each list rebuild.  One 16-byte
    // record per slot: the integrator, which reads them by slot in particle order, gathers four
    // records instead of twelve scattered floats.
    float4* fpair = nullptr;   // cutoff pairs
    float4* fintra = nullptr;  // intra / bonded-compound pairs
    float4* fb = nullptr;      // bonded terms
    float4* fpme = nullptr;    // PME gather
    float* ephi = nullptr;     // [slots] 1/2 q phi of charged slots (energy evaluations), summed in slot
                               // order by the energy reduction (the gather's bin order is not fixed)
    double* acc = nullptr;     // [ACC_N]
    double* acc_part = nullptr;// [REDUCE_BLOCKS][ACC_N] first-stage sums of the energy reduction
    double* epart = nullptr;   // [epart_cap * ACC_N] per-wave energy records
    int64_t epart_cap = 0;
    ERegions ereg;
    float* fmax = nullptr;     // [1]
    float* thermo = nullptr;   // [2] velocity scale factor applied by the integrator, last measured temperature [K]
    bool temp_measured = false;

    hipStream_t s_main = nullptr, s_pme = nullptr;   // both at the default priority; s_pme: k_local, then PME
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // one MD step (evaluation without energies + integration) as an instantiated HIP graph; its
    // kernel arguments depend on the list, so every list rebuild invalidates it
    hipGraphExec_t step_graph = nullptr;
    bool step_graph_valid = false;
    int64_t graph_captures = 0;

    int64_t step = 0;
    int64_t steps_since_build = 0;
    int n_rebuilds = 0;
    lima_energy energy{};
    int64_t energy_step = -1;      // MD step of the last energy evaluation issued (-1: none, -2: not an MD step)
    bool energy_unread = false;    // its totals are still only on the device (e.acc)
    bool forces_valid = false;
    bool packed = false;       // the integrator already packed the slots for the next evaluation
    bool critical = false;

    // timing: per step 7 events (before pack, after pack, pair, intra, bonded, pme, integrate),
    // recorded into a ring without host syncs and summed when the ring is flushed
    bool timing = false;
    static constexpr int TRING = 256, TEV = 7;
    hipEvent_t tev[TRING * TEV];
    int tpos = 0;
    lima_timings tacc{};
    int64_t pairs_last = 0;      // pairs inside the cutoff at the last energy evaluation
    double host_issue_ms = 0;    // host time spent issuing MD steps (excluding the final device wait)
    int64_t host_steps = 0;

    // data log: a ring of log_cap frames written at every step % data_logging_interval == 0.  A
    // frame holds, per entry of LIMA's buffer layout (compound c slot s at c*32 + s, tinymol t at
    // nC*32 + t), (position [nm], potE) and (force, speed), plus the step's FP64 accumulators.
    int log_cap = 0;
    int64_t log_P = 0;             // entries per frame, nC*32 + nT
    float4* log_pe = nullptr;      // [log_cap][log_P]
    float4* log_fv = nullptr;      // [log_cap][log_P]
    double* log_acc = nullptr;     // [log_cap][ACC_N]
    int* log_index = nullptr;      // [N] buffer entry of each particle
    int* log_key = nullptr;        // [N] global index of the particle's compound key particle, -1 for tinymols
    int64_t log_frames = 0;        // frames written since create
    std::vector<int64_t> log_step; // [log_cap] step of the frame held in each ring slot
};

#define HIP_CHECK(x)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) throw ::lima::HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern thread_local std::string g_err;   // lima_last_error() text, set by every failing C-ABI call

struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };
struct LimaFail : std::runtime_error {
    lima_status code;
    LimaFail(lima_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---- nlist.hip ----
void nlist_alloc(Engine& e);
void nlist_free(Engine& e);
// Rebuilds clusters + lists from positions w.  checked: wait for the build, read the status words,
// grow a pair list that overflowed and build again (the one host synchronisation of a rebuild);
// otherwise the words are copied to pinned host memory behind an event and examined later
// (poll_rebuild_flags / settle in engine.hip).
void list_interval_check(Engine& e, hipStream_t s);   // raises DF_STALE if the open interval's list may miss a pair
void nlist_build(Engine& e, const uint4* w, bool checked);
void nlist_grow(Engine& e, int need);   // pair-list capacity >= 1.5 x need (the step graph is captured again)
void reset_flags(Engine& e, int64_t snapshots, hipStream_t s);   // clear the error words, set DF_SNAPSHOTS

// ---- forces.hip ----
int pair_split(const Engine& e);
ERegions energy_regions(const Engine& e);
void launch_reduce_energy(Engine& e, bool kinetic, hipStream_t s);
void launch_pack(Engine& e, const uint4* w, hipStream_t s);
BinArgs bin_args(const Engine& e);   // PME tile bins (bin_count == nullptr when PME is not tiled)
inline int slots_max(const Engine& e) { return e.nl.ncl_max * CLUSTER; }   // length of the slot arrays
// 32-bit Coord differences are exact when the box, the list radius and a column's width stay below
// 2^31 Coord units together (boxes up to ~9.4 nm); the list and pair kernels pick their path by it
inline bool narrow_coords(const Engine& e) {
    const long long R = llround(((double)e.rc + (double)e.skin) * 1e8);
    const long long Lmax = std::max(e.Lq[0], std::max(e.Lq[1], e.Lq[2]));
    return 2 * Lmax + 2 * R < (1ll << 31);
}
int nlist_count(Engine& e);          // current cluster count read back from the device (synchronises)
void launch_pair(Engine& e, bool energy, hipStream_t s);
void launch_intra(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s);
void launch_bonded(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s);
void launch_local(Engine& e, const uint4* w, const int4* img, bool energy, hipStream_t s);   // intra + bonded, one launch

// ---- pme.hip ----
void pme_alloc(Engine& e, const std::vector<float>& charges_global, const std::vector<uint4>& w);
void pme_free(Engine& e);
void pme_grow_bins(Engine& e);       // tile-bin capacity x 2 (the step graph is captured again)
int pme_zpass_lines(const Pme& m);             // kx lines per workgroup of the PME z pass
int pme_virial_records(const Engine& e);       // energy records of the z pass (one per workgroup)
void pme_clear_bins(Engine& e, hipStream_t s);
void launch_pme(Engine& e, const uint4* w, bool energy, hipStream_t s);

// ---- scan helper (nlist.hip) ----
void launch_exclusive_scan(const int* in, int* out, int n, hipStream_t s);

}  // namespace lima
