// lima_gen.cpp -- deterministic synthetic box generators (see include/lima_gen.h).
//
// Molecule -> compound partitioning follows LIMA's builder semantics: a molecule is cut into
// compounds of <= 32 consecutive particles (Constants.h:91), every pair of particles that share
// a bond, angle, dihedral or improper is marked in the BondedParticlesLUT of the two compounds
// (CompoundBuilder.cpp:17-60), self pairs are always marked (CompoundBuilder.cpp:62-68), and a
// compound's origo is the integer-nm node of its key particle (LimaPositionSystem.cuh:112-124).
// Parameters are stored in LIMA units: charges e -> kC/mol, masses g -> kg, epsilon and force
// constants kJ -> J (Forcefield.cpp:342-347, Bodies.cpp:9-47) and dihedral k halved
// (DihedralBond::Parameters::CreateFromCharmm, Bodies.cpp:36-38).
#include "../../include/lima_gen.h"
#include "lima_box_impl.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr float PI_L = 3.14159f;                                  // Constants.h:51
constexpr float DEG_TO_RAD_L = PI_L / 180.f;                      // Constants.h:57
constexpr double AVOGADRO = 6.02214076e23;
constexpr double BOLTZMANN = (double)1.38066e-23f;                // Constants.h:59
constexpr float E_TO_KCMOL = (float)(1.602176634e-19 * AVOGADRO / 1000.0);   // Constants.h:65
constexpr double NANO_TO_LIMA = 1e8;
constexpr double GMOL_PER_NM3_AT_1GCC = 602.214076;               // 1 g/cm^3 in (g/mol)/nm^3

// ---- RNG: SplitMix64 seeding a xoshiro256** stream ----
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) { uint64_t x = seed; for (auto& v : s) v = splitmix(x); }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }      // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {   // Box-Muller
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    uint64_t below(uint64_t n) { return next() % n; }
};

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double k) const { return {x * k, y * k, z * k}; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    double len() const { return std::sqrt(dot(*this)); }
    V3 norm() const { const double l = len(); return {x / l, y / l, z / l}; }
};

struct Atom { V3 pos; double q_e, sigma, eps_kj, mass_g; };
struct Bond { int a, b; double b0, kb_kj; };
struct Angle { int a, b, c; double th0_deg, k_kj, ub0, kub_kj; };
struct Dihedral { int a, b, c, d; double phi0_deg, k_kj; int n; };
struct Molecule {
    std::vector<Atom> atoms;
    std::vector<Bond> bonds;
    std::vector<Angle> angles;
    std::vector<Dihedral> dihedrals;
};
struct Tiny { V3 pos; double sigma, eps_kj, mass_g; };

// ---- water model (config 0 / TIP3P-like parameters, resources/forcefields/charmm27.ff/ffnonbonded.itp:132-133) ----
constexpr double W_BOND = 0.09572, W_ANGLE_DEG = 104.52, W_KB = 345000.0, W_KA = 383.0;
constexpr double O_Q = -0.834, H_Q = 0.417, O_SIG = 0.315061, O_EPS = 0.636386, O_M = 15.9994, H_M = 1.008;
constexpr double WATER_MASS = O_M + 2 * H_M;

V3 rotate(const double q[4], const V3& v) {   // q = (w, x, y, z) unit quaternion
    const V3 u(q[1], q[2], q[3]);
    const double s = q[0];
    const V3 t = u.cross(v) * 2.0;
    return v + t * s + u.cross(t);
}

void random_quat(Rng& r, double q[4]) {   // Shoemake 1992, uniform on SO(3)
    const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    q[0] = b * std::cos(2 * M_PI * u3); q[1] = a * std::sin(2 * M_PI * u2);
    q[2] = a * std::cos(2 * M_PI * u2); q[3] = b * std::sin(2 * M_PI * u3);
}

Molecule make_water(const V3& o, Rng& r) {
    double q[4];
    random_quat(r, q);
    const double half = 0.5 * W_ANGLE_DEG * M_PI / 180.0;
    const V3 h1(W_BOND * std::sin(half), W_BOND * std::cos(half), 0), h2(-W_BOND * std::sin(half), W_BOND * std::cos(half), 0);
    Molecule m;
    m.atoms.push_back({o, O_Q, O_SIG, O_EPS, O_M});
    m.atoms.push_back({o + rotate(q, h1), H_Q, 0.0, 0.0, H_M});
    m.atoms.push_back({o + rotate(q, h2), H_Q, 0.0, 0.0, H_M});
    m.bonds.push_back({0, 1, W_BOND, W_KB});
    m.bonds.push_back({0, 2, W_BOND, W_KB});
    m.angles.push_back({1, 0, 2, W_ANGLE_DEG, W_KA, 0.0, 0.0});
    return m;
}

// ---- chain model (config 1) ----
constexpr double C_BOND = 0.15, C_KB = 2e5, C_ANGLE = 111.0, C_KA = 400.0, C_M = 14.027;

void chain_topology(Molecule& m, int n, Rng& r) {
    for (int i = 0; i + 1 < n; i++) m.bonds.push_back({i, i + 1, C_BOND, C_KB});
    for (int i = 0; i + 2 < n; i++) m.angles.push_back({i, i + 1, i + 2, C_ANGLE, C_KA, 0.0, 0.0});
    for (int i = 0; i + 3 < n; i++) m.dihedrals.push_back({i, i + 1, i + 2, i + 3, 0.0, r.uniform(0.5, 5.0), 3});
}

void chain_params(Molecule& m, int n, Rng& r) {
    double qs = 0;
    for (int i = 0; i < n; i++) {
        Atom a{};
        a.sigma = r.uniform(0.30, 0.38);
        a.eps_kj = r.uniform(0.2, 0.9);
        a.q_e = r.uniform(-0.5, 0.5);
        a.mass_g = C_M;
        qs += a.q_e;
        m.atoms.push_back(a);
    }
    for (auto& a : m.atoms) a.q_e -= qs / n;   // neutralise per chain
}

// place d from (a, b, c) with |cd| = bond, angle bcd, dihedral abcd (NeRF)
V3 nerf(const V3& a, const V3& b, const V3& c, double bond, double ang, double dih) {
    const V3 bc = (c - b).norm();
    V3 n = (b - a).cross(bc);
    if (n.len() < 1e-12) n = V3(1, 0, 0).cross(bc);
    if (n.len() < 1e-12) n = V3(0, 1, 0).cross(bc);
    n = n.norm();
    const V3 m = n.cross(bc);
    const V3 d2(-bond * std::cos(ang), bond * std::sin(ang) * std::cos(dih), bond * std::sin(ang) * std::sin(dih));
    return c + bc * d2.x + m * d2.y + n * d2.z;
}

// Spatial hash with periodic minimum image, used for clash tests during placement.
struct Hash {
    double L[3]; double cell; int n[3];
    std::unordered_map<int64_t, std::vector<V3>> map;
    Hash(const double* box, double c) : cell(c) {
        for (int d = 0; d < 3; d++) { L[d] = box[d]; n[d] = std::max(1, (int)std::floor(box[d] / c)); }
    }
    int64_t key(int x, int y, int z) const { return ((int64_t)z * n[1] + y) * n[0] + x; }
    void idx(const V3& p, int* c) const {
        const double v[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; d++) {
            double w = std::fmod(v[d], L[d]); if (w < 0) w += L[d];
            c[d] = std::min(n[d] - 1, (int)(w / L[d] * n[d]));
        }
    }
    void add(const V3& p) { int c[3]; idx(p, c); map[key(c[0], c[1], c[2])].push_back(p); }
    double mindist(const V3& p, double cutoff) const {
        int c[3]; idx(p, c);
        double best = 1e30;
        const int r[3] = {std::min(n[0] / 2, (int)std::ceil(cutoff / (L[0] / n[0]))), std::min(n[1] / 2, (int)std::ceil(cutoff / (L[1] / n[1]))),
                          std::min(n[2] / 2, (int)std::ceil(cutoff / (L[2] / n[2])))};
        std::vector<int64_t> seen;
        for (int dz = -r[2]; dz <= r[2]; dz++) for (int dy = -r[1]; dy <= r[1]; dy++) for (int dx = -r[0]; dx <= r[0]; dx++) {
            const int64_t k = key(((c[0] + dx) % n[0] + n[0]) % n[0], ((c[1] + dy) % n[1] + n[1]) % n[1], ((c[2] + dz) % n[2] + n[2]) % n[2]);
            if (std::find(seen.begin(), seen.end(), k) != seen.end()) continue;
            seen.push_back(k);
            auto it = map.find(k);
            if (it == map.end()) continue;
            for (const V3& q : it->second) {
                double dd[3] = {q.x - p.x, q.y - p.y, q.z - p.z};
                for (int d = 0; d < 3; d++) dd[d] -= L[d] * std::round(dd[d] / L[d]);
                best = std::min(best, std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]));
            }
        }
        return best;
    }
};

double mind_img(const V3& a, const V3& b, const double* L) {
    double dd[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
    for (int d = 0; d < 3; d++) dd[d] -= L[d] * std::round(dd[d] / L[d]);
    return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
}

// Self-avoiding chain growth with trans/gauche dihedrals and backtracking.  `accept`
// decides whether a candidate position is allowed given the atoms placed so far.
template <class Accept>
bool grow_chain(std::vector<V3>& pos, int n, const V3& p0, const V3& dir, Rng& r, double trans_bias, double jitter_deg, Accept accept) {
    const double ang = C_ANGLE * M_PI / 180.0;
    pos.clear();
    pos.push_back(p0);
    pos.push_back(p0 + dir.norm() * C_BOND);
    // third atom: angle at atom 1, random azimuth
    int budget = 20000;
    std::vector<int> fails(n + 1, 0);
    while ((int)pos.size() < n && budget-- > 0) {
        const int k = (int)pos.size();
        bool placed = false;
        for (int attempt = 0; attempt < 12 && !placed; attempt++) {
            V3 cand;
            if (k == 2) {
                const V3 ref = pos[0] + V3(r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1));
                cand = nerf(ref, pos[0], pos[1], C_BOND, ang, r.uniform(-M_PI, M_PI));
            } else {
                const double u = r.uniform();
                double dih = u < trans_bias ? 180.0 : (u < 0.5 * (1 + trans_bias) ? 60.0 : -60.0);
                dih += r.uniform(-jitter_deg, jitter_deg);
                cand = nerf(pos[k - 3], pos[k - 2], pos[k - 1], C_BOND, ang, dih * M_PI / 180.0);
            }
            if (accept(cand, pos)) { pos.push_back(cand); placed = true; }
        }
        if (!placed) {
            if (++fails[k] > 8 && k > 3) {   // backtrack
                const int back = std::min(k - 3, 1 + (int)r.below(4));
                for (int b = 0; b < back; b++) { fails[pos.size()] = 0; pos.pop_back(); }
            }
        }
    }
    return (int)pos.size() == n;
}

// ---- box assembly: molecules -> LIMA compounds + LUTs ----
struct Builder {
    double box[3];
    std::vector<Molecule> mols;
    std::vector<Tiny> tinies;
};

}  // namespace

namespace {

int64_t fdiv(int64_t a, int64_t b) { int64_t q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) --q; return q; }

lima_box* assemble(const Builder& B, Rng& r, double temperature) {
    auto* box = new lima_box();
    int64_t Lq[3];
    for (int d = 0; d < 3; d++) { Lq[d] = llround(B.box[d] * NANO_TO_LIMA); box->box[d] = (double)Lq[d] / NANO_TO_LIMA; }
    // compounds
    struct CRef { int mol, first, n; };
    std::vector<CRef> comps;
    std::vector<std::vector<int>> mol_atom_comp(B.mols.size()), mol_atom_slot(B.mols.size());
    for (size_t mi = 0; mi < B.mols.size(); mi++) {
        const int na = (int)B.mols[mi].atoms.size();
        mol_atom_comp[mi].resize(na); mol_atom_slot[mi].resize(na);
        for (int f = 0; f < na; f += 32) {
            const int n = std::min(32, na - f);
            for (int k = 0; k < n; k++) { mol_atom_comp[mi][f + k] = (int)comps.size(); mol_atom_slot[mi][f + k] = k; }
            comps.push_back({(int)mi, f, n});
        }
    }
    const int nC = (int)comps.size();
    box->nC = nC;
    box->cn.resize(nC); box->origo.resize(nC * 3); box->key.resize(nC);
    box->coord.assign((size_t)nC * 32 * 3, 0); box->vel.assign((size_t)nC * 32 * 3, 0.f);
    box->q.assign((size_t)nC * 32, 0.f); box->sig.assign((size_t)nC * 32, 0.f); box->eps.assign((size_t)nC * 32, 0.f);
    box->mass.assign((size_t)nC * 32, 0.f);
    box->lut.assign((size_t)nC * 5 * 32, 0u);
    // velocities: Maxwell-Boltzmann, then remove the centre-of-mass drift
    struct Pv { double v[3]; double m; };
    std::vector<Pv> pv;
    for (int c = 0; c < nC; c++) {
        const auto& cr = comps[c];
        const Molecule& m = B.mols[cr.mol];
        box->cn[c] = cr.n;
        V3 cen;
        for (int k = 0; k < cr.n; k++) cen = cen + m.atoms[cr.first + k].pos;
        cen = cen * (1.0 / cr.n);
        int key = 0; double best = 1e30;
        for (int k = 0; k < cr.n; k++) { const double d = (m.atoms[cr.first + k].pos - cen).len(); if (d < best) { best = d; key = k; } }
        box->key[c] = key;
        const V3& kp = m.atoms[cr.first + key].pos;
        const double kv[3] = {kp.x, kp.y, kp.z};
        int64_t shift[3], org[3];
        for (int d = 0; d < 3; d++) {
            const int64_t ku = llround(kv[d] * NANO_TO_LIMA);
            const int64_t kw = ku - fdiv(ku, Lq[d]) * Lq[d];
            shift[d] = kw - ku;
            org[d] = fdiv(kw, (int64_t)NANO_TO_LIMA);
            box->origo[c * 3 + d] = (int32_t)org[d];
        }
        for (int k = 0; k < cr.n; k++) {
            const Atom& a = m.atoms[cr.first + k];
            const double pv3[3] = {a.pos.x, a.pos.y, a.pos.z};
            const size_t ps = (size_t)c * 32 + k;
            for (int d = 0; d < 3; d++) box->coord[ps * 3 + d] = (int32_t)(llround(pv3[d] * NANO_TO_LIMA) + shift[d] - org[d] * (int64_t)NANO_TO_LIMA);
            box->q[ps] = (float)a.q_e * E_TO_KCMOL;
            box->sig[ps] = (float)a.sigma;
            box->eps[ps] = (float)a.eps_kj * 1000.f;
            box->mass[ps] = (float)(a.mass_g / 1000.0);
            Pv p{}; p.m = a.mass_g / 1000.0;
            const double s = temperature > 0 ? std::sqrt(BOLTZMANN * AVOGADRO * temperature / p.m) : 0.0;
            for (int d = 0; d < 3; d++) p.v[d] = s * r.normal();
            pv.push_back(p);
        }
    }
    box->nT = (int)B.tinies.size();
    for (const Tiny& t : B.tinies) {
        Pv p{}; p.m = t.mass_g / 1000.0;
        const double s = temperature > 0 ? std::sqrt(BOLTZMANN * AVOGADRO * temperature / p.m) : 0.0;
        for (int d = 0; d < 3; d++) p.v[d] = s * r.normal();
        pv.push_back(p);
    }
    if (temperature > 0 && pv.size() > 1) {
        double mt = 0, mv[3] = {0, 0, 0};
        for (auto& p : pv) { mt += p.m; for (int d = 0; d < 3; d++) mv[d] += p.m * p.v[d]; }
        for (auto& p : pv) for (int d = 0; d < 3; d++) p.v[d] -= mv[d] / mt;
    }
    size_t pk = 0;
    for (int c = 0; c < nC; c++)
        for (int k = 0; k < comps[c].n; k++, pk++)
            for (int d = 0; d < 3; d++) box->vel[((size_t)c * 32 + k) * 3 + d] = (float)pv[pk].v[d];
    for (const Tiny& t : B.tinies) {
        const double tv[3] = {t.pos.x, t.pos.y, t.pos.z};
        for (int d = 0; d < 3; d++) {
            int64_t u = llround(tv[d] * NANO_TO_LIMA);
            u -= fdiv(u, Lq[d]) * Lq[d];
            const int64_t blk = fdiv(u, (int64_t)NANO_TO_LIMA);
            box->tblock.push_back((int32_t)blk);
            box->tcoord.push_back((int32_t)(u - blk * (int64_t)NANO_TO_LIMA));
            box->tvel.push_back((float)pv[pk].v[d]);
        }
        pk++;
        box->tsig.push_back((float)t.sigma); box->teps.push_back((float)t.eps_kj * 1000.f); box->tmass.push_back((float)(t.mass_g / 1000.0));
    }
    // exclusion LUTs and bonded terms
    auto gid = [&](size_t mi, int a) { return mol_atom_comp[mi][a] * 32 + mol_atom_slot[mi][a]; };
    auto mark = [&](size_t mi, std::initializer_list<int> ids) {
        for (int a : ids) for (int b : ids) {
            if (a == b) continue;
            const int ca = mol_atom_comp[mi][a], cb = mol_atom_comp[mi][b];
            if (std::abs(ca - cb) > 2) throw std::runtime_error("bonded term spans compounds more than 2 apart");
            box->lut[((size_t)ca * 5 + (cb - ca + 2)) * 32 + mol_atom_slot[mi][a]] |= 1u << mol_atom_slot[mi][b];
        }
    };
    for (size_t mi = 0; mi < B.mols.size(); mi++) {
        const Molecule& m = B.mols[mi];
        for (auto& b : m.bonds) {
            mark(mi, {b.a, b.b});
            box->bond_ids.push_back(gid(mi, b.a)); box->bond_ids.push_back(gid(mi, b.b));
            box->bond_p.push_back((float)b.b0); box->bond_p.push_back((float)b.kb_kj * 1000.f);
        }
        for (auto& a : m.angles) {
            mark(mi, {a.a, a.b, a.c});
            box->angle_ids.push_back(gid(mi, a.a)); box->angle_ids.push_back(gid(mi, a.b)); box->angle_ids.push_back(gid(mi, a.c));
            box->angle_p.push_back((float)a.th0_deg * DEG_TO_RAD_L); box->angle_p.push_back((float)a.k_kj * 1000.f);
            box->angle_p.push_back((float)a.ub0); box->angle_p.push_back((float)a.kub_kj * 1000.f);
        }
        for (auto& d : m.dihedrals) {
            mark(mi, {d.a, d.b, d.c, d.d});
            box->dih_ids.push_back(gid(mi, d.a)); box->dih_ids.push_back(gid(mi, d.b));
            box->dih_ids.push_back(gid(mi, d.c)); box->dih_ids.push_back(gid(mi, d.d));
            box->dih_p.push_back((float)d.phi0_deg * DEG_TO_RAD_L); box->dih_p.push_back((float)d.k_kj * 1000.f / 2.f);
            box->dih_p.push_back((float)d.n);
        }
    }
    for (int c = 0; c < nC; c++)
        for (int s = 0; s < 32; s++) box->lut[((size_t)c * 5 + 2) * 32 + s] |= 1u << s;   // no self interaction
    return box;
}

// Places up to n_target waters on a shuffled lattice of the given spacing, skipping sites whose
// O or H would come closer than o_min / h_min to the solute; returns the number placed.
int fill_waters(Builder& B, Rng& r, int n_target, const Hash& solute, double o_min, double h_min, double spacing) {
    int n[3];
    for (int d = 0; d < 3; d++) n[d] = std::max(1, (int)std::lround(B.box[d] / spacing));
    std::vector<int64_t> sites((size_t)n[0] * n[1] * n[2]);
    for (size_t k = 0; k < sites.size(); k++) sites[k] = (int64_t)k;
    for (size_t k = sites.size(); k > 1; k--) std::swap(sites[k - 1], sites[r.below(k)]);
    int placed = 0;
    for (int64_t s : sites) {
        if (placed >= n_target) break;
        const int x = (int)(s % n[0]), y = (int)((s / n[0]) % n[1]), z = (int)(s / ((int64_t)n[0] * n[1]));
        const V3 o((x + 0.5) * B.box[0] / n[0], (y + 0.5) * B.box[1] / n[1], (z + 0.5) * B.box[2] / n[2]);
        if (!solute.map.empty() && solute.mindist(o, o_min) < o_min) continue;
        Molecule w = make_water(o, r);
        if (!solute.map.empty() && (solute.mindist(w.atoms[1].pos, h_min) < h_min || solute.mindist(w.atoms[2].pos, h_min) < h_min)) continue;
        B.mols.push_back(std::move(w));
        placed++;
    }
    return placed;
}

double total_mass_g(int n_chain_atoms, int n_water) { return n_chain_atoms * C_M + n_water * WATER_MASS; }

}  // namespace

extern "C" {

const char* lima_gen_last_error(void) { return g_err.c_str(); }

lima_status lima_gen_water_lattice(int32_t npd, double L, uint64_t seed, double T, lima_box** out) {
    try {
        Rng r(seed);
        Builder B{};
        for (int d = 0; d < 3; d++) B.box[d] = L;
        const double a = L / npd;
        for (int z = 0; z < npd; z++) for (int y = 0; y < npd; y++) for (int x = 0; x < npd; x++)
            B.mols.push_back(make_water(V3((x + 0.5) * a, (y + 0.5) * a, (z + 0.5) * a), r));
        *out = assemble(B, r, T);
        return LIMA_OK;
    } catch (const std::exception& e) { g_err = e.what(); return LIMA_ERR_BAD_INPUT; }
}

lima_status lima_gen_chain_water(double bx, double by, double bz, int32_t n_chains, int32_t len, int32_t n_waters,
                                 uint64_t seed, double T, lima_box** out) {
    try {
        Rng r(seed);
        Builder B{};
        B.box[0] = bx; B.box[1] = by; B.box[2] = bz;
        Hash solute(B.box, 0.5);
        // Two facing leaflets of chains along z: leaflet A points +z, leaflet B points -z.  The
        // chains sit on a hexagonal lattice of 0.5 nm spacing (rows of 8) centred in x-y,
        // leaflet A in the first half of the rows, B in the second.
        const int per_row = 8;
        const int rows = (n_chains + per_row - 1) / per_row;
        const double sp = 0.5, rsp = sp * std::sqrt(3.0) / 2.0;
        const double x0 = 0.5 * (bx - (per_row - 0.5) * sp), y0 = 0.5 * (by - (rows - 1) * rsp);
        for (int c = 0; c < n_chains; c++) {
            const int row = c / per_row, col = c % per_row;
            const bool up = row < rows / 2 || rows == 1;
            V3 p0(x0 + col * sp + (row % 2) * 0.5 * sp + r.uniform(-0.03, 0.03), y0 + row * rsp + r.uniform(-0.03, 0.03),
                  up ? 0.35 + r.uniform(0, 0.1) : bz - 0.35 - r.uniform(0, 0.1));
            // all-trans zig-zag whose axis is z: rise b sin(theta/2) per atom, alternating lateral
            // offset +-b cos(theta/2)/2 along a random azimuth, plus a small positional jitter
            Molecule m;
            chain_params(m, len, r);
            chain_topology(m, len, r);
            const double half = 0.5 * C_ANGLE * M_PI / 180.0;
            const double rise = C_BOND * std::sin(half), lat = 0.5 * C_BOND * std::cos(half);
            const double az = r.uniform(0, 2 * M_PI);
            const V3 u(std::cos(az), std::sin(az), 0);
            std::vector<V3> pos(len);
            for (int k = 0; k < len; k++) {
                const double sgn = (k % 2) ? 1.0 : -1.0;
                pos[k] = p0 + V3(0, 0, (up ? 1.0 : -1.0) * rise * k) + u * (sgn * lat) +
                         V3(r.uniform(-0.005, 0.005), r.uniform(-0.005, 0.005), r.uniform(-0.005, 0.005));
            }
            for (int k = 0; k < len; k++) { m.atoms[k].pos = pos[k]; }
            for (auto& p : pos) solute.add(p);
            B.mols.push_back(std::move(m));
        }
        // Waters on a lattice at 1 g/cm^3 first; when the free volume next to the chains cannot hold
        // n_waters there, the lattice is made finer and the clash distances shorter, pass by pass
        // (the structure is relaxed by minimisation before any dynamics).  Fewer than n_waters
        // after the last pass is an error, never a silently smaller box.
        const size_t before = B.mols.size();
        double omin = 0.30, hmin = 0.22, wsp = std::cbrt(WATER_MASS / GMOL_PER_NM3_AT_1GCC);
        int placed = 0;
        for (int pass = 0; pass < 12; pass++) {
            B.mols.resize(before);
            placed = fill_waters(B, r, n_waters, solute, omin, hmin, wsp);
            if (placed >= n_waters) break;
            omin *= 0.97; hmin *= 0.97; wsp *= 0.97;
        }
        if (placed < n_waters)
            throw std::runtime_error("chain + water box: placed " + std::to_string(placed) + " of " + std::to_string(n_waters) +
                                     " waters; enlarge the box or request fewer");
        *out = assemble(B, r, T);
        return LIMA_OK;
    } catch (const std::exception& e) { g_err = e.what(); return LIMA_ERR_BAD_INPUT; }
}

lima_status lima_gen_em_system(int32_t target, uint64_t seed, double noise, lima_box** out) {
    try {
        Rng r(seed);
        const int clen = 50;
        int nch = (int)std::lround(0.15 * target / clen);
        if (nch < 1) nch = 1;
        int nw = (int)std::lround((target - nch * clen) / 3.0);
        if (nw < 0) nw = 0;
        const double L = std::cbrt(total_mass_g(nch * clen, nw) / GMOL_PER_NM3_AT_1GCC);
        Builder B{};
        for (int d = 0; d < 3; d++) B.box[d] = L;
        Hash solute(B.box, std::min(0.5, L / 3));
        const double L3[3] = {L, L, L};
        for (int c = 0; c < nch; c++) {
            Molecule m;
            chain_params(m, clen, r);
            chain_topology(m, clen, r);
            std::vector<V3> pos;
            bool ok = false;
            for (int tries = 0; tries < 50 && !ok; tries++) {
                const V3 p0(r.uniform(0, L), r.uniform(0, L), r.uniform(0, L));
                const V3 dir(r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1));
                auto accept = [&](const V3& cand, const std::vector<V3>& cur) {
                    for (size_t j = 0; j + 3 < cur.size(); j++) if (mind_img(cand, cur[j], L3) < 0.3) return false;
                    return solute.map.empty() || solute.mindist(cand, 0.3) >= 0.3;
                };
                ok = grow_chain(pos, clen, p0, dir.len() > 0 ? dir : V3(0, 0, 1), r, 0.5, 20.0, accept);
            }
            if (!ok) throw std::runtime_error("em chain placement failed");
            for (int k = 0; k < clen; k++) m.atoms[k].pos = pos[k];
            for (auto& p : pos) solute.add(p);
            B.mols.push_back(std::move(m));
        }
        // waters: relax the clash distances until every water is placed
        const size_t before = B.mols.size();
        double omin = 0.30, hmin = 0.22, sp = std::cbrt(WATER_MASS / GMOL_PER_NM3_AT_1GCC) * 0.92;
        for (int pass = 0; pass < 8 && (int)(B.mols.size() - before) < nw; pass++) {
            B.mols.resize(before);
            fill_waters(B, r, nw, solute, omin, hmin, sp);
            omin *= 0.85; hmin *= 0.85; sp *= 0.95;
        }
        // perturb every coordinate by Gaussian noise to create clashes
        for (auto& m : B.mols) for (auto& a : m.atoms) a.pos = a.pos + V3(r.normal(), r.normal(), r.normal()) * noise;
        *out = assemble(B, r, 0.0);
        return LIMA_OK;
    } catch (const std::exception& e) { g_err = e.what(); return LIMA_ERR_BAD_INPUT; }
}

lima_status lima_gen_large_water(int32_t nw, int32_t nx, double xs, double xe, double xm, uint64_t seed, double T, lima_box** out) {
    try {
        Rng r(seed);
        const double L = std::cbrt((nw * WATER_MASS + nx * xm) / GMOL_PER_NM3_AT_1GCC);
        Builder B{};
        for (int d = 0; d < 3; d++) B.box[d] = L;
        int n = (int)std::ceil(std::cbrt((double)(nw + nx)));
        const double a = L / n;
        std::vector<int64_t> sites((size_t)n * n * n);
        for (size_t k = 0; k < sites.size(); k++) sites[k] = (int64_t)k;
        for (size_t k = sites.size(); k > 1; k--) std::swap(sites[k - 1], sites[r.below(k)]);
        for (int k = 0; k < nw + nx; k++) {
            const int64_t s = sites[k];
            const V3 p((s % n + 0.5) * a, ((s / n) % n + 0.5) * a, (s / ((int64_t)n * n) + 0.5) * a);
            if (k < nw) B.mols.push_back(make_water(p, r));
            else B.tinies.push_back({p, xs, xe, xm});
        }
        *out = assemble(B, r, T);
        return LIMA_OK;
    } catch (const std::exception& e) { g_err = e.what(); return LIMA_ERR_BAD_INPUT; }
}

void lima_box_describe(const lima_box* b, lima_box_desc* d) {
    std::memset(d, 0, sizeof(*d));
    for (int k = 0; k < 3; k++) d->box_nm[k] = b->box[k];
    d->n_compounds = b->nC; d->n_tinymols = b->nT;
    d->compound_n_particles = b->cn.data(); d->compound_origo = b->origo.data(); d->compound_key_particle = b->key.data();
    d->particle_coord = b->coord.data(); d->particle_velocity = b->vel.data(); d->particle_charge = b->q.data();
    d->particle_sigma = b->sig.data(); d->particle_epsilon = b->eps.data(); d->particle_mass = b->mass.data();
    d->bonded_lut = b->lut.data();
    d->tinymol_block = b->tblock.data(); d->tinymol_coord = b->tcoord.data(); d->tinymol_velocity = b->tvel.data();
    d->tinymol_sigma = b->tsig.data(); d->tinymol_epsilon = b->teps.data(); d->tinymol_mass = b->tmass.data();
    d->n_bonds = (int32_t)b->bond_p.size() / 2; d->bond_ids = b->bond_ids.data(); d->bond_params = b->bond_p.data();
    d->n_angles = (int32_t)b->angle_p.size() / 4; d->angle_ids = b->angle_ids.data(); d->angle_params = b->angle_p.data();
    d->n_dihedrals = (int32_t)b->dih_p.size() / 3; d->dihedral_ids = b->dih_ids.data(); d->dihedral_params = b->dih_p.data();
    d->n_impropers = (int32_t)b->imp_p.size() / 2; d->improper_ids = b->imp_ids.data(); d->improper_params = b->imp_p.data();
    d->n_pairs = (int32_t)b->pair_p.size() / 2; d->pair_ids = b->pair_ids.data(); d->pair_params = b->pair_p.data();
}

void lima_box_free(lima_box* b) { delete b; }

int64_t lima_box_n_particles(const lima_box* b) {
    int64_t n = b->nT;
    for (int c = 0; c < b->nC; c++) n += b->cn[c];
    return n;
}

uint64_t lima_box_hash(const lima_box* b) {   // FNV-1a over the integer / float payload
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    mix(b->box, sizeof(b->box));
    mix(b->cn.data(), b->cn.size() * 4); mix(b->origo.data(), b->origo.size() * 4); mix(b->coord.data(), b->coord.size() * 4);
    mix(b->vel.data(), b->vel.size() * 4); mix(b->q.data(), b->q.size() * 4); mix(b->sig.data(), b->sig.size() * 4);
    mix(b->eps.data(), b->eps.size() * 4); mix(b->lut.data(), b->lut.size() * 4);
    mix(b->tcoord.data(), b->tcoord.size() * 4); mix(b->bond_p.data(), b->bond_p.size() * 4); mix(b->dih_p.data(), b->dih_p.size() * 4);
    return h;
}

void lima_default_params(lima_simparams* p, float dt_fs, int64_t n_steps) {
    std::memset(p, 0, sizeof(*p));
    p->n_steps = n_steps;
    p->dt_fs = dt_fs;
    p->em_force_tolerance = 100.f;
    p->bc_select = LIMA_BC_PBC;
    p->enable_electrostatics = 1;
    p->cutoff_nm = 1.2f;
    p->data_logging_interval = 5;
    p->steps_per_temperature_measurement = 200;
    p->pme_spacing_nm = 0.12f;
    p->nlist_skin_nm = 0.30f;   // lists stay complete for ~10-15 steps of 2 fs water (see nlist_interval)
    p->nlist_interval = 40;
    p->energy_interval = 0;
}

}  // extern "C"
