// lima_topo.cpp -- .gro / .top / .itp inputs to a LIMA box: readers, forcefield parameter lookup
// and the molecule -> compound builder.  Interface and the reference behaviour: include/lima_topo.h.
#include "../../include/lima_topo.h"
#include "lima_box_impl.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <optional>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct TopoFail : std::runtime_error {
    lima_status code;
    TopoFail(lima_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(const std::string& m) { throw TopoFail(LIMA_ERR_BAD_INPUT, m); }

// LIMA's constants (Constants.h:43-66) and parameter conversions (Bodies.cpp:9-47)
constexpr float KILO = 1000.f;
constexpr float PI_L = 3.14159f;
constexpr float DEG_TO_RAD_L = PI_L / 180.f;
constexpr float E_TO_KCMOL = (float)(1.602176634e-19 * 6.02214076e23 / 1000.0);
constexpr int64_t NANO_TO_LIMA = 100000000;
constexpr int MAXP = LIMA_MAX_COMPOUND_PARTICLES;

struct BondP { float b0, kb; };                      // nm, J/mol/nm^2
struct PairP { float sigma, eps; };                  // nm, J/mol
struct AngleP { float th0, kth, ub0, kub; };         // rad, J/mol/rad^2, nm, J/mol/nm^2
struct DihP { float phi0, kphi, n; };                // rad, J/mol (GROMACS k * 1000 / 2), multiplicity
struct ImpP { float psi0, kpsi; };                   // rad, J/mol/rad^2
bool operator==(const BondP& a, const BondP& b) { return a.b0 == b.b0 && a.kb == b.kb; }
bool operator==(const PairP& a, const PairP& b) { return a.sigma == b.sigma && a.eps == b.eps; }
bool operator==(const AngleP& a, const AngleP& b) { return a.th0 == b.th0 && a.kth == b.kth && a.ub0 == b.ub0 && a.kub == b.kub; }
bool operator==(const DihP& a, const DihP& b) { return a.phi0 == b.phi0 && a.kphi == b.kphi && a.n == b.n; }
bool operator==(const ImpP& a, const ImpP& b) { return a.psi0 == b.psi0 && a.kpsi == b.kpsi; }

BondP bond_from_gmx(float b0, float kb) { return {b0, kb * KILO}; }
PairP pair_from_gmx(float s, float e) { return {s, e * KILO}; }
AngleP angle_from_gmx(float t0, float k, float ub0, float kub) { return {t0 * DEG_TO_RAD_L, k * KILO, ub0, kub * KILO}; }
DihP dih_from_gmx(float p0, float k, int n) { return {p0 * DEG_TO_RAD_L, k * KILO / 2.f, (float)n}; }
ImpP imp_from_gmx(float p0, float k) { return {p0 * DEG_TO_RAD_L, k * KILO}; }

// ---- text helpers ----
std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
std::vector<std::string> words(const std::string& s) {
    std::istringstream is(s);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    return w;
}
int to_int(const std::string& s, const std::string& where) {
    try {
        return std::stoi(s);
    } catch (...) {
        fail("expected an integer, got '" + s + "' in " + where);
    }
}
double to_num(const std::string& s, const std::string& where) {
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) fail("expected a number, got '" + s + "' in " + where);
    return v;
}
std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
bool exists(const std::string& p) {
    std::ifstream f(p);
    return f.good();
}
bool ends_with(const std::string& s, const std::string& e) { return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0; }

// ---- .gro (GroFile, MDFiles.cpp:13-55,82-160): fixed columns ----
struct GroAtom {
    int resnr = 0, id = 0;
    std::string resname, name;
    double x[3]{};
    bool has_v = false;
    double v[3]{};
};
struct Gro {
    std::string title;
    std::vector<GroAtom> atoms;
    double box[3]{};
};

Gro read_gro(const std::string& path) {
    if (!ends_with(path, ".gro")) fail("expected .gro extension: " + path);
    std::ifstream f(path);
    if (!f) fail("file \"" + path + "\" was not found");
    Gro g;
    std::string line;
    if (!std::getline(f, g.title)) fail("empty .gro file " + path);
    if (!g.title.empty() && g.title.back() == '\r') g.title.pop_back();
    if (!std::getline(f, line)) fail("missing atom count in " + path);
    const int n = to_int(trim(line), path + " (atom count)");
    if (n < 0) fail("negative atom count in " + path);
    g.atoms.resize(n);
    for (int i = 0; i < n; i++) {
        if (!std::getline(f, line)) fail(path + ": expected " + std::to_string(n) + " atom lines, found " + std::to_string(i));
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const std::string where = path + " line " + std::to_string(i + 3);
        if (line.size() < 44) fail(where + " is shorter than the 44 fixed columns of a .gro atom");
        GroAtom& a = g.atoms[i];
        a.resnr = to_int(line.substr(0, 5), where);
        a.resname = trim(line.substr(5, 5));
        a.name = trim(line.substr(10, 5));
        a.id = to_int(line.substr(15, 5), where);
        for (int d = 0; d < 3; d++) a.x[d] = to_num(line.substr(20 + 8 * d, 8), where);
        if (line.size() >= 68) {
            a.has_v = true;
            for (int d = 0; d < 3; d++) a.v[d] = to_num(line.substr(44 + 8 * d, 8), where);
        }
    }
    if (!std::getline(f, line)) fail("missing box line in " + path);
    const auto w = words(line);
    if (w.size() < 3) fail("box line of " + path + " holds fewer than 3 numbers");
    for (int d = 0; d < 3; d++) g.box[d] = to_num(w[d], path + " (box)");
    for (size_t k = 3; k < w.size(); k++)
        if (to_num(w[k], path + " (box)") != 0.0) throw TopoFail(LIMA_ERR_UNSUPPORTED, "only orthorhombic boxes are supported: " + path);
    return g;
}

// ---- topology (TopologyFile, Topologyfile.cpp:8-400) ----
template <int N, class P>
struct Term {
    std::array<int, N> ids{};
    std::optional<P> p;
};
struct AtomEntry {
    std::string type, residue, name;
    int resnr = 0;
    float charge_e = 0.f;
    float mass_g = NAN;   // NaN: not given, taken from the atom type
};
struct Mtype {
    std::string name;
    std::vector<AtomEntry> atoms;
    std::map<int, int> nr_to_local;
    std::vector<Term<2, BondP>> bonds;
    std::vector<Term<2, PairP>> pairs;
    std::vector<Term<3, AngleP>> angles;
    std::vector<Term<4, DihP>> dihs;
    std::vector<Term<4, ImpP>> imps;
};

enum Sec { S_NONE, S_MOLTYPE, S_ATOMS, S_BONDS, S_PAIRS, S_ANGLES, S_DIHS, S_IMPS, S_SYSTEM, S_MOLECULES,
           S_ATOMTYPES, S_PAIRTYPES, S_BONDTYPES, S_ANGLETYPES, S_DIHTYPES, S_IMPTYPES, S_DEFAULTS, S_IGNORE, S_COUNT };

struct Topo {
    std::map<std::string, std::shared_ptr<Mtype>> mtypes;
    std::shared_ptr<Mtype> current;
    bool has_system = false;
    std::string system_title;
    std::vector<std::shared_ptr<Mtype>> molecules;   // one entry per molecule instance
    std::vector<std::string> ff[S_COUNT];             // raw forcefield lines per section
    bool has_forcefield = false;
    std::set<std::string> defines{"FLEXIBLE"};
    std::vector<std::string> include_dirs;
};

// Section of a [ directive ]; per file, the second "dihedrals" / "dihedraltypes" are impropers.
Sec section_of(const std::string& d, int& n_dih, int& n_dihtypes, const std::string& where) {
    if (d == "moleculetype") return S_MOLTYPE;
    if (d == "atoms") return S_ATOMS;
    if (d == "bonds") return S_BONDS;
    if (d == "pairs") return S_PAIRS;
    if (d == "angles") return S_ANGLES;
    if (d == "dihedrals") {
        if (++n_dih > 2) fail("more than 2 [ dihedrals ] sections in one file: " + where);
        return n_dih == 1 ? S_DIHS : S_IMPS;
    }
    if (d == "system") return S_SYSTEM;
    if (d == "molecules") return S_MOLECULES;
    if (d == "atomtypes") return S_ATOMTYPES;
    if (d == "pairtypes") return S_PAIRTYPES;
    if (d == "bondtypes") return S_BONDTYPES;
    if (d == "angletypes") return S_ANGLETYPES;
    if (d == "dihedraltypes") {
        if (++n_dihtypes > 2) fail("more than 2 [ dihedraltypes ] sections in one file: " + where);
        return n_dihtypes == 1 ? S_DIHTYPES : S_IMPTYPES;
    }
    if (d == "impropertypes") return S_IMPTYPES;
    if (d == "defaults") return S_DEFAULTS;
    if (d == "position_restraints" || d == "cmap" || d == "cmaptypes" || d == "constrainttypes" || d == "settles" ||
        d == "exclusions" || d == "nonbond_params")
        return S_IGNORE;
    fail("unexpected topology directive [ " + d + " ] in " + where);
}

template <int N>
bool map_ids(const Mtype& m, const std::vector<std::string>& w, std::array<int, N>& ids, const std::string& where) {
    for (int i = 0; i < N; i++) {
        const int nr = to_int(w[i], where);
        const auto it = m.nr_to_local.find(nr);
        if (it == m.nr_to_local.end()) return false;   // LIMA skips terms naming unknown atoms
        ids[i] = it->second;
    }
    return true;
}

void parse_entry(Mtype& m, Sec s, const std::vector<std::string>& w, const std::string& where) {
    auto need = [&](size_t n) { if (w.size() < n) fail("too few fields in " + where); };
    auto f = [&](size_t k) { return (float)to_num(w[k], where); };
    switch (s) {
    case S_ATOMS: {
        need(7);
        AtomEntry a;
        const int nr = to_int(w[0], where);
        a.type = w[1];
        a.resnr = to_int(w[2], where);
        a.residue = w[3];
        a.name = w[4];
        a.charge_e = f(6);
        if (w.size() >= 8) a.mass_g = f(7);
        m.nr_to_local[nr] = (int)m.atoms.size();
        m.atoms.push_back(a);
        break;
    }
    case S_BONDS: {
        need(3);
        Term<2, BondP> t;
        if (!map_ids<2>(m, w, t.ids, where)) break;
        if (w.size() >= 5) t.p = bond_from_gmx(f(3), f(4));
        m.bonds.push_back(t);
        break;
    }
    case S_PAIRS: {
        need(3);
        Term<2, PairP> t;
        if (!map_ids<2>(m, w, t.ids, where)) break;
        if (w.size() >= 5) t.p = pair_from_gmx(f(3), f(4));
        m.pairs.push_back(t);
        break;
    }
    case S_ANGLES: {   // funct 5: theta0 k ub0 kub (Urey-Bradley); otherwise theta0 k [B-state values]
        need(4);
        Term<3, AngleP> t;
        if (!map_ids<3>(m, w, t.ids, where)) break;
        const int funct = to_int(w[3], where);
        if (funct == 5 && w.size() >= 8) t.p = angle_from_gmx(f(4), f(5), f(6), f(7));
        else if (funct != 5 && w.size() >= 6) t.p = angle_from_gmx(f(4), f(5), 0.f, 0.f);
        m.angles.push_back(t);
        break;
    }
    case S_DIHS: {
        need(5);
        Term<4, DihP> t;
        if (!map_ids<4>(m, w, t.ids, where)) break;
        if (w.size() >= 8) t.p = dih_from_gmx(f(5), f(6), to_int(w[7], where));
        m.dihs.push_back(t);
        break;
    }
    case S_IMPS: {
        need(5);
        Term<4, ImpP> t;
        if (!map_ids<4>(m, w, t.ids, where)) break;
        if (w.size() >= 7) t.p = imp_from_gmx(f(5), f(6));
        m.imps.push_back(t);
        break;
    }
    default:
        break;
    }
}

void parse_topology(Topo& T, const std::string& path, int depth) {
    if (depth > 32) fail("#include nesting deeper than 32 at " + path);
    std::ifstream file(path);
    if (!file) fail("failed to open " + path);
    Sec cur = S_NONE;
    int n_dih = 0, n_dihtypes = 0;
    std::vector<std::pair<bool, bool>> cond;   // (this branch active, an earlier branch was taken)
    auto active = [&] {
        for (auto& c : cond) if (!c.first) return false;
        return true;
    };
    std::string raw;
    int lineno = 0;
    while (std::getline(file, raw)) {
        lineno++;
        const std::string where = path + ":" + std::to_string(lineno);
        std::string t = trim(raw);
        if (!t.empty() && t[0] == '#') {
            const auto w = words(t);
            const std::string d = w[0];
            if (d == "#ifdef" || d == "#ifndef") {
                if (w.size() < 2) fail("missing name after " + d + " at " + where);
                const bool def = T.defines.count(w[1]) != 0;
                const bool take = d == "#ifdef" ? def : !def;
                cond.push_back({take, take});
            } else if (d == "#else") {
                if (cond.empty()) fail("#else without #ifdef at " + where);
                cond.back().first = !cond.back().second;
                cond.back().second = true;
            } else if (d == "#endif") {
                if (cond.empty()) fail("#endif without #ifdef at " + where);
                cond.pop_back();
            } else if (!active()) {
                continue;
            } else if (d == "#define") {
                if (w.size() >= 2) T.defines.insert(w[1]);
            } else if (d == "#undef") {
                if (w.size() >= 2) T.defines.erase(w[1]);
            } else if (d == "#include") {
                if (w.size() < 2 || w[1].size() < 3) fail("include is not formatted as expected at " + where);
                const std::string name = w[1].substr(1, w[1].size() - 2);
                if (name.find("posre") != std::string::npos) continue;   // position restraints: not supported
                std::string found;
                if (name[0] == '/') { if (exists(name)) found = name; }
                else if (exists(dir_of(path) + "/" + name)) found = dir_of(path) + "/" + name;
                for (size_t k = 0; found.empty() && k < T.include_dirs.size(); k++)
                    if (exists(T.include_dirs[k] + "/" + name)) found = T.include_dirs[k] + "/" + name;
                if (found.empty()) fail("could not find include \"" + name + "\" of " + where);
                parse_topology(T, found, depth + 1);
            }
            continue;
        }
        if (!active()) continue;
        const size_t sc = t.find(';');
        if (sc != std::string::npos) t = trim(t.substr(0, sc));
        if (t.empty()) continue;
        if (t[0] == '[') {
            const size_t e = t.find(']');
            if (e == std::string::npos) fail("unterminated directive at " + where);
            cur = section_of(trim(t.substr(1, e - 1)), n_dih, n_dihtypes, where);
            if (cur == S_DEFAULTS) {
                if (T.has_forcefield) fail("a topology can include one forcefield; a second [ defaults ] at " + where);
                T.has_forcefield = true;
            }
            continue;
        }
        const auto w = words(t);
        switch (cur) {
        case S_MOLTYPE: {
            auto m = std::make_shared<Mtype>();
            m->name = w[0];
            if (T.mtypes.count(m->name)) fail("moleculetype " + m->name + " defined twice (" + where + ")");
            T.mtypes[m->name] = m;
            T.current = m;
            break;
        }
        case S_ATOMS: case S_BONDS: case S_PAIRS: case S_ANGLES: case S_DIHS: case S_IMPS:
            if (!T.current) fail("[ atoms ] or a bonded section before any [ moleculetype ] at " + where);
            parse_entry(*T.current, cur, w, where);
            break;
        case S_SYSTEM:
            if (!T.has_system) { T.has_system = true; T.system_title = t; }
            break;
        case S_MOLECULES: {
            if (!T.has_system) fail("[ molecules ] before [ system ] at " + where);
            if (w.size() < 2) fail("too few fields in " + where);
            const auto it = T.mtypes.find(w[0]);
            if (it == T.mtypes.end()) fail("moleculetype " + w[0] + " not defined before being used at " + where);
            const int cnt = to_int(w[1], where);
            for (int k = 0; k < cnt; k++) T.molecules.push_back(it->second);
            break;
        }
        case S_ATOMTYPES: case S_PAIRTYPES: case S_BONDTYPES: case S_ANGLETYPES: case S_DIHTYPES: case S_IMPTYPES:
            T.ff[cur].push_back(t);
            break;
        default:
            break;
        }
    }
    if (!cond.empty()) fail("#ifdef without #endif in " + path);
}

// ---- forcefield (LIMAForcefield, Forcefield.cpp:41-479) ----
struct AtomType {
    float mass_kg = 0.f, charge = 0.f, sigma = 0.f, eps = 0.f;
};

template <int N, class P>
struct TypeDB {
    struct E {
        std::array<std::string, N> t;
        P p;
    };
    std::vector<E> entries;
    static std::array<std::string, N> flip(std::array<std::string, N> t) {
        if constexpr (N == 2) std::swap(t[0], t[1]);
        if constexpr (N == 3) std::swap(t[0], t[2]);
        return t;
    }
    void add(const std::array<std::string, N>& t, const P& p) {
        for (const E& e : entries)
            if (e.t == t && e.p == p) return;   // duplicates in a forcefield are ignored
        entries.push_back({t, p});
    }
    // both orders are stored (for 4-atom types the caller passes its own flip)
    void insert(const std::array<std::string, N>& t, const std::array<std::string, N>& flipped, const P& p) {
        add(t, p);
        if (flipped != t) add(flipped, p);
    }
    // every entry matching q with the fewest "X" wildcards
    std::vector<P> match(const std::array<std::string, N>& q) const {
        std::vector<P> best;
        int lowest = 1 << 30;
        for (const E& e : entries) {
            int wild = 0;
            bool ok = true;
            for (int i = 0; i < N && ok; i++) {
                if (e.t[i] == q[i]) continue;
                if (e.t[i] == "X") { wild++; continue; }
                ok = false;
            }
            if (!ok) continue;
            if (wild < lowest) { lowest = wild; best.clear(); }
            if (wild == lowest) best.push_back(e.p);
        }
        return best;
    }
};

struct Forcefield {
    std::map<std::string, AtomType> atypes;
    TypeDB<2, BondP> bonds;
    TypeDB<2, PairP> pairs;
    TypeDB<3, AngleP> angles;
    TypeDB<4, DihP> dihs;
    TypeDB<4, ImpP> imps;

    const AtomType& atom(const std::string& t) const {
        const auto it = atypes.find(t);
        if (it == atypes.end()) fail("failed to find atomtype [" + t + "]");
        return it->second;
    }
};

Forcefield load_forcefield(const Topo& T) {
    Forcefield F;
    for (const std::string& l : T.ff[S_ATOMTYPES]) {
        const auto w = words(l);
        // name [bond_type] [at.num] mass charge ptype sigma epsilon: located from the ptype column
        size_t k = 0;
        for (size_t i = 3; i < w.size(); i++)
            if (w[i] == "A" || w[i] == "D" || w[i] == "S" || w[i] == "V") { k = i; break; }
        if (k == 0 || k + 2 >= w.size()) fail("cannot read atomtype line: " + l);
        AtomType a;
        a.mass_kg = (float)to_num(w[k - 2], l) / KILO;
        a.charge = (float)to_num(w[k - 1], l) * E_TO_KCMOL;
        a.sigma = (float)to_num(w[k + 1], l);
        a.eps = (float)to_num(w[k + 2], l) * KILO;
        if (!F.atypes.count(w[0])) F.atypes[w[0]] = a;   // the first definition wins
    }
    auto f = [](const std::vector<std::string>& w, size_t k, const std::string& l) {
        if (k >= w.size()) fail("too few fields in forcefield line: " + l);
        return (float)to_num(w[k], l);
    };
    for (const std::string& l : T.ff[S_BONDTYPES]) {
        const auto w = words(l);
        const std::array<std::string, 2> t{w.at(0), w.at(1)};
        F.bonds.insert(t, TypeDB<2, BondP>::flip(t), bond_from_gmx(f(w, 3, l), f(w, 4, l)));
    }
    for (const std::string& l : T.ff[S_PAIRTYPES]) {
        const auto w = words(l);
        const std::array<std::string, 2> t{w.at(0), w.at(1)};
        F.pairs.insert(t, TypeDB<2, PairP>::flip(t), pair_from_gmx(f(w, 3, l), f(w, 4, l)));
    }
    for (const std::string& l : T.ff[S_ANGLETYPES]) {
        const auto w = words(l);
        const std::array<std::string, 3> t{w.at(0), w.at(1), w.at(2)};
        const bool ub = w.size() > 3 && w[3] == "5";   // funct 5: Urey-Bradley terms follow
        const float ub0 = ub ? f(w, 6, l) : 0.f, kub = ub ? f(w, 7, l) : 0.f;
        F.angles.insert(t, TypeDB<3, AngleP>::flip(t), angle_from_gmx(f(w, 4, l), f(w, 5, l), ub0, kub));
    }
    for (const std::string& l : T.ff[S_DIHTYPES]) {
        const auto w = words(l);
        const std::array<std::string, 4> t{w.at(0), w.at(1), w.at(2), w.at(3)};
        const std::array<std::string, 4> r{t[3], t[2], t[1], t[0]};
        F.dihs.insert(t, r, dih_from_gmx(f(w, 5, l), f(w, 6, l), (int)f(w, 7, l)));
    }
    for (const std::string& l : T.ff[S_IMPTYPES]) {
        const auto w = words(l);
        const std::array<std::string, 4> t{w.at(0), w.at(1), w.at(2), w.at(3)};
        const std::array<std::string, 4> r{t[3], t[1], t[2], t[0]};   // LIMA swaps only the outer atoms
        F.imps.insert(t, r, imp_from_gmx(f(w, 5, l), f(w, 6, l)));
    }
    return F;
}

template <int N>
std::string type_list(const std::array<std::string, N>& t) {
    std::string s;
    for (int i = 0; i < N; i++) s += (i ? " " : "") + t[i];
    return s;
}

// ---- the built system ----
struct GTerm {   // a term on global particle indices
    std::array<int, 4> ids{};
    std::array<float, 4> p{};
};

}  // namespace

struct lima_system {
    Gro gro;
    std::vector<int> layout_gro;            // [nC*32 + nT] gro atom, -1 for unused slots
    std::vector<std::vector<int>> tiny_atoms;   // gro atoms of each tinymol's molecule
    int n_molecules = 0;
    lima_box box;
};

namespace {

int64_t fdiv(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

// Splits one molecule (atoms in ascending order, bonds among them) into groups of at most 32.
// The spanning tree is rooted at an end of a longest path; a node's side branches of at most 32
// atoms are placed whole (smallest first, a new group when one does not fit), longer branches
// atom by atom, and the heaviest branch last, so a long chain continues in the next group.
std::vector<std::vector<int>> partition_molecule(const std::vector<int>& atoms, const std::vector<std::vector<int>>& adj) {
    const int n = (int)atoms.size();
    if (n <= MAXP) return {atoms};
    std::map<int, int> local;
    for (int i = 0; i < n; i++) local[atoms[i]] = i;
    std::vector<std::vector<int>> nb(n);
    for (int i = 0; i < n; i++) {
        for (int g : adj[atoms[i]]) nb[i].push_back(local.at(g));
        std::sort(nb[i].begin(), nb[i].end());
    }
    auto bfs = [&](int s, std::vector<int>& order, std::vector<int>& parent) {
        order.clear();
        parent.assign(n, -2);
        parent[s] = -1;
        order.push_back(s);
        for (size_t h = 0; h < order.size(); h++)
            for (int v : nb[order[h]])
                if (parent[v] == -2) { parent[v] = order[h]; order.push_back(v); }
    };
    std::vector<int> order, parent;
    bfs(0, order, parent);
    const int root = order.back();   // farthest from the first atom: an end of a longest path
    bfs(root, order, parent);
    std::vector<std::vector<int>> kids(n);
    for (int v : order)
        if (parent[v] >= 0) kids[parent[v]].push_back(v);
    std::vector<int> size(n, 1);
    for (int k = n - 1; k > 0; k--) size[parent[order[k]]] += size[order[k]];

    std::vector<std::vector<int>> groups(1);
    auto place = [&](int v) {
        if ((int)groups.back().size() == MAXP) groups.emplace_back();
        groups.back().push_back(atoms[v]);
    };
    auto place_subtree = [&](int v) {   // pre-order, fits in the current group
        std::vector<int> st{v};
        while (!st.empty()) {
            const int u = st.back();
            st.pop_back();
            groups.back().push_back(atoms[u]);
            for (auto it = kids[u].rbegin(); it != kids[u].rend(); ++it) st.push_back(*it);
        }
    };
    // stack of (node, whole subtree)
    std::vector<std::pair<int, bool>> st{{root, false}};
    while (!st.empty()) {
        const auto [v, whole] = st.back();
        st.pop_back();
        if (whole) {
            if (size[v] > MAXP - (int)groups.back().size()) groups.emplace_back();
            place_subtree(v);
            continue;
        }
        place(v);
        std::vector<int> c = kids[v];
        std::stable_sort(c.begin(), c.end(), [&](int a, int b) { return size[a] < size[b]; });
        for (auto it = c.rbegin(); it != c.rend(); ++it) st.push_back({*it, size[*it] <= MAXP});
    }
    if (groups.back().empty()) groups.pop_back();
    return groups;
}

lima_system* build(const char* gro_path, const char* top_path, const char* const* dirs, int n_dirs, int flags) {
    if (!gro_path || !top_path) fail("null path");
    if (flags & ~LIMA_TOPO_KEEP_SMALL_MOLECULES) fail("unknown flags");
    auto S = std::make_unique<lima_system>();
    S->gro = read_gro(gro_path);
    Topo T;
    for (int k = 0; k < n_dirs; k++)
        if (dirs && dirs[k]) T.include_dirs.push_back(dirs[k]);
    if (!ends_with(top_path, ".top") && !ends_with(top_path, ".itp")) fail("expected a .top or .itp topology: " + std::string(top_path));
    parse_topology(T, top_path, 0);
    if (!T.has_system) fail("the topology has no [ system ]");
    if (!T.has_forcefield) fail("the topology includes no forcefield ([ defaults ])");
    const Forcefield F = load_forcefield(T);

    // particles in [ molecules ] order = .gro order
    struct Part { const AtomEntry* a; };
    std::vector<Part> parts;
    std::vector<GTerm> bonds, pairs, angles, dihs, imps;
    for (const auto& m : T.molecules) {
        const int off = (int)parts.size();
        for (const AtomEntry& a : m->atoms) parts.push_back({&a});
        auto types = [&](auto ids) {
            std::array<std::string, std::tuple_size<decltype(ids)>::value> t;
            for (size_t i = 0; i < t.size(); i++) t[i] = parts[off + ids[i]].a->type;
            return t;
        };
        auto gid = [&](auto ids) {
            std::array<int, 4> g{-1, -1, -1, -1};
            for (size_t i = 0; i < ids.size(); i++) g[i] = off + ids[i];
            return g;
        };
        for (const auto& t : m->bonds) {
            std::vector<BondP> ps;
            if (t.p) ps.push_back(*t.p);
            else ps = F.bonds.match(types(t.ids));
            if (ps.empty()) fail("failed to find bond parameters for types " + type_list<2>(types(t.ids)));
            for (const auto& p : ps) bonds.push_back({gid(t.ids), {p.b0, p.kb, 0.f, 0.f}});
        }
        for (const auto& t : m->pairs) {
            std::vector<PairP> ps;
            if (t.p) ps.push_back(*t.p);
            else ps = F.pairs.match(types(t.ids));
            if (ps.empty()) {   // combined from the atom types (Forcefield.cpp:431-446)
                const AtomType& l = F.atom(types(t.ids)[0]);
                const AtomType& r = F.atom(types(t.ids)[1]);
                ps.push_back({(l.sigma + r.sigma) * 0.5f, std::sqrt(l.eps * r.eps)});
            }
            for (const auto& p : ps) pairs.push_back({gid(t.ids), {p.sigma, p.eps, 0.f, 0.f}});
        }
        for (const auto& t : m->angles) {
            std::vector<AngleP> ps;
            if (t.p) ps.push_back(*t.p);
            else ps = F.angles.match(types(t.ids));
            if (ps.empty()) fail("failed to find angle parameters for types " + type_list<3>(types(t.ids)));
            for (const auto& p : ps) angles.push_back({gid(t.ids), {p.th0, p.kth, p.ub0, p.kub}});
        }
        for (const auto& t : m->dihs) {
            std::vector<DihP> ps;
            if (t.p) ps.push_back(*t.p);
            else ps = F.dihs.match(types(t.ids));
            if (ps.empty()) fail("failed to find dihedral parameters for types " + type_list<4>(types(t.ids)));
            for (const auto& p : ps) dihs.push_back({gid(t.ids), {p.phi0, p.kphi, p.n, 0.f}});
        }
        for (const auto& t : m->imps) {
            std::vector<ImpP> ps;
            if (t.p) ps.push_back(*t.p);
            else ps = F.imps.match(types(t.ids));
            if (ps.empty()) fail("failed to find improper parameters for types " + type_list<4>(types(t.ids)));
            for (const auto& p : ps) imps.push_back({gid(t.ids), {p.psi0, p.kpsi, 0.f, 0.f}});
        }
    }
    const int N = (int)parts.size();
    if (N != (int)S->gro.atoms.size())
        fail("the topology's molecules hold " + std::to_string(N) + " atoms, the .gro " + std::to_string(S->gro.atoms.size()));
    for (const Part& p : parts) F.atom(p.a->type);   // every atom type must exist

    // molecules: connected components of the bond graph, ordered by their first atom
    std::vector<std::vector<int>> adj(N);
    for (const GTerm& b : bonds) {
        adj[b.ids[0]].push_back(b.ids[1]);
        adj[b.ids[1]].push_back(b.ids[0]);
    }
    std::vector<int> comp(N, -1);
    std::vector<std::vector<int>> mols;
    for (int i = 0; i < N; i++) {
        if (comp[i] >= 0) continue;
        std::vector<int> q{i};
        comp[i] = (int)mols.size();
        for (size_t h = 0; h < q.size(); h++)
            for (int v : adj[q[h]])
                if (comp[v] < 0) { comp[v] = comp[i]; q.push_back(v); }
        std::sort(q.begin(), q.end());
        mols.push_back(q);
    }
    const bool keep_small = (flags & LIMA_TOPO_KEEP_SMALL_MOLECULES) != 0;
    std::vector<std::vector<int>> groups;   // compounds, in order
    std::vector<std::vector<int>> tinies;
    for (const auto& m : mols) {
        const bool lxx = parts[m[0]].a->residue == "lxx";
        if (m.size() > 3 || lxx || keep_small) {
            S->n_molecules++;
            for (auto& g : partition_molecule(m, adj)) groups.push_back(std::move(g));
        } else {
            tinies.push_back(m);
        }
    }
    const int nC = (int)groups.size(), nT = (int)tinies.size();
    std::vector<int> p_comp(N, -1), p_slot(N, -1);
    for (int c = 0; c < nC; c++)
        for (int s = 0; s < (int)groups[c].size(); s++) { p_comp[groups[c][s]] = c; p_slot[groups[c][s]] = s; }

    // box arrays
    lima_box& B = S->box;
    int64_t Lq[3];
    for (int d = 0; d < 3; d++) {
        Lq[d] = llround(S->gro.box[d] * (double)NANO_TO_LIMA);
        if (Lq[d] <= 0) fail("box edges must be positive");
        B.box[d] = (double)Lq[d] / (double)NANO_TO_LIMA;
    }
    B.nC = nC;
    B.nT = nT;
    B.cn.resize(nC);
    B.origo.resize((size_t)nC * 3);
    B.key.resize(nC);
    B.coord.assign((size_t)nC * MAXP * 3, 0);
    B.vel.assign((size_t)nC * MAXP * 3, 0.f);
    B.q.assign((size_t)nC * MAXP, 0.f);
    B.sig.assign((size_t)nC * MAXP, 0.f);
    B.eps.assign((size_t)nC * MAXP, 0.f);
    B.mass.assign((size_t)nC * MAXP, 0.f);
    B.lut.assign((size_t)nC * 5 * MAXP, 0u);
    S->layout_gro.assign((size_t)nC * MAXP + nT, -1);
    auto qpos = [&](int g, int d) { return llround(S->gro.atoms[g].x[d] * (double)NANO_TO_LIMA); };
    for (int c = 0; c < nC; c++) {
        const auto& gr = groups[c];
        const int n = (int)gr.size();
        B.cn[c] = n;
        // positions made whole relative to the compound's first particle (minimum image)
        std::vector<std::array<int64_t, 3>> u(n);
        for (int s = 0; s < n; s++)
            for (int d = 0; d < 3; d++) {
                int64_t x = qpos(gr[s], d);
                if (s > 0) {
                    const int64_t dx = x - u[0][d];
                    x -= llround((double)dx / (double)Lq[d]) * Lq[d];
                }
                u[s][d] = x;
            }
        double cen[3] = {0, 0, 0};
        for (int s = 0; s < n; s++)
            for (int d = 0; d < 3; d++) cen[d] += (double)u[s][d] / n;
        int key = 0;
        double best = 1e300;
        for (int s = 0; s < n; s++) {
            double r2 = 0;
            for (int d = 0; d < 3; d++) r2 += ((double)u[s][d] - cen[d]) * ((double)u[s][d] - cen[d]);
            if (r2 < best) { best = r2; key = s; }
        }
        B.key[c] = key;
        for (int d = 0; d < 3; d++) {
            const int64_t kw = u[key][d] - fdiv(u[key][d], Lq[d]) * Lq[d];   // key particle wrapped into the box
            const int64_t shift = kw - u[key][d];
            const int64_t org = fdiv(kw, NANO_TO_LIMA);
            B.origo[(size_t)c * 3 + d] = (int32_t)org;
            for (int s = 0; s < n; s++) {
                const int64_t rel = u[s][d] + shift - org * NANO_TO_LIMA;
                if (rel < INT32_MIN || rel > INT32_MAX) fail("a compound is too large for int32 Coord");
                B.coord[((size_t)c * MAXP + s) * 3 + d] = (int32_t)rel;
            }
        }
        for (int s = 0; s < n; s++) {
            const int g = gr[s];
            const AtomEntry& a = *parts[g].a;
            const AtomType& at = F.atom(a.type);
            const size_t ps = (size_t)c * MAXP + s;
            B.q[ps] = a.charge_e * E_TO_KCMOL;
            B.mass[ps] = std::isnan(a.mass_g) ? at.mass_kg : a.mass_g / KILO;
            B.sig[ps] = at.sigma;
            B.eps[ps] = at.eps;
            if (S->gro.atoms[g].has_v)
                for (int d = 0; d < 3; d++) B.vel[ps * 3 + d] = (float)(S->gro.atoms[g].v[d] * 1000.0);   // nm/ps -> nm/ns
            S->layout_gro[ps] = g;
        }
    }
    for (int t = 0; t < nT; t++) {   // single site at the first atom, its type's LJ and mass
        const int g = tinies[t][0];
        const std::string& type = parts[g].a->type;
        const AtomType& at = F.atom(type);
        float m = at.mass_kg;
        if (type == "OW" && F.atypes.count("HW")) m += 2.f * F.atypes.at("HW").mass_kg;
        for (int d = 0; d < 3; d++) {
            int64_t x = qpos(g, d);
            x -= fdiv(x, Lq[d]) * Lq[d];
            const int64_t blk = fdiv(x, NANO_TO_LIMA);
            B.tblock.push_back((int32_t)blk);
            B.tcoord.push_back((int32_t)(x - blk * NANO_TO_LIMA));
            B.tvel.push_back(S->gro.atoms[g].has_v ? (float)(S->gro.atoms[g].v[d] * 1000.0) : 0.f);
        }
        B.tsig.push_back(at.sigma);
        B.teps.push_back(at.eps);
        B.tmass.push_back(m);
        S->layout_gro[(size_t)nC * MAXP + t] = g;
        S->tiny_atoms.push_back(tinies[t]);
    }
    // terms of compound particles only (a tinymol's terms are dropped); exclusions from every
    // bond, angle, dihedral and improper
    auto in_compounds = [&](const GTerm& t, int n) {
        for (int i = 0; i < n; i++) if (p_comp[t.ids[i]] < 0) return false;
        return true;
    };
    auto lid = [&](int g) { return p_comp[g] * MAXP + p_slot[g]; };
    auto exclude = [&](const GTerm& t, int n) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                const int a = t.ids[i], b = t.ids[j];
                const int ca = p_comp[a], cb = p_comp[b];
                if (std::abs(ca - cb) > 2)
                    throw TopoFail(LIMA_ERR_UNSUPPORTED, "a bonded term spans compounds " + std::to_string(ca) + " and " +
                                                             std::to_string(cb) + "; exclusion tables reach 2 compounds");
                B.lut[((size_t)ca * 5 + (cb - ca + 2)) * MAXP + p_slot[a]] |= 1u << p_slot[b];
            }
    };
    for (const GTerm& t : bonds)
        if (in_compounds(t, 2)) {
            exclude(t, 2);
            B.bond_ids.push_back(lid(t.ids[0])); B.bond_ids.push_back(lid(t.ids[1]));
            B.bond_p.push_back(t.p[0]); B.bond_p.push_back(t.p[1]);
        }
    for (const GTerm& t : angles)
        if (in_compounds(t, 3)) {
            exclude(t, 3);
            for (int i = 0; i < 3; i++) B.angle_ids.push_back(lid(t.ids[i]));
            for (int i = 0; i < 4; i++) B.angle_p.push_back(t.p[i]);
        }
    for (const GTerm& t : dihs)
        if (in_compounds(t, 4)) {
            exclude(t, 4);
            for (int i = 0; i < 4; i++) B.dih_ids.push_back(lid(t.ids[i]));
            for (int i = 0; i < 3; i++) B.dih_p.push_back(t.p[i]);
        }
    for (const GTerm& t : imps)
        if (in_compounds(t, 4)) {
            exclude(t, 4);
            for (int i = 0; i < 4; i++) B.imp_ids.push_back(lid(t.ids[i]));
            for (int i = 0; i < 2; i++) B.imp_p.push_back(t.p[i]);
        }
    for (const GTerm& t : pairs)
        if (in_compounds(t, 2)) {
            B.pair_ids.push_back(lid(t.ids[0])); B.pair_ids.push_back(lid(t.ids[1]));
            B.pair_p.push_back(t.p[0]); B.pair_p.push_back(t.p[1]);
        }
    for (int c = 0; c < nC; c++)
        for (int s = 0; s < MAXP; s++) B.lut[((size_t)c * 5 + 2) * MAXP + s] |= 1u << s;   // no self interaction
    return S.release();
}

template <class F>
lima_status guarded(F&& f) {
    try {
        f();
        return LIMA_OK;
    } catch (const TopoFail& x) {
        g_err = x.what();
        return x.code;
    } catch (const std::exception& x) {
        g_err = x.what();
        return LIMA_ERR_BAD_INPUT;
    }
}

// composeGroLine (MDFiles.cpp:57-80)
void write_gro_line(std::ostream& o, const GroAtom& a, const double x[3]) {
    o << std::setw(5) << std::left << a.resnr << std::setw(5) << std::left << a.resname << std::setw(5) << std::left << a.name
      << std::setw(5) << std::right << a.id;
    for (int d = 0; d < 3; d++) o << std::setw(8) << std::fixed << std::setprecision(3) << (float)x[d];
    if (a.has_v)
        for (int d = 0; d < 3; d++) o << std::setw(8) << std::fixed << std::setprecision(4) << (float)a.v[d];
    o << "\n";
}

}  // namespace

extern "C" {

const char* lima_topo_last_error(void) { return g_err.c_str(); }

lima_status lima_load_system(const char* gro_path, const char* top_path, const char* const* include_dirs, int32_t n_include_dirs,
                             int32_t flags, lima_system** out) {
    if (!out) { g_err = "null output"; return LIMA_ERR_BAD_INPUT; }
    *out = nullptr;
    return guarded([&] { *out = build(gro_path, top_path, include_dirs, n_include_dirs, flags); });
}

void lima_system_free(lima_system* sys) { delete sys; }

lima_status lima_system_get_info(const lima_system* s, lima_system_info* out) {
    if (!s || !out) { g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    out->n_gro_atoms = (int64_t)s->gro.atoms.size();
    out->n_molecules = s->n_molecules;
    out->n_compounds = s->box.nC;
    out->n_tinymols = s->box.nT;
    out->n_bonds = (int32_t)(s->box.bond_p.size() / 2);
    out->n_angles = (int32_t)(s->box.angle_p.size() / 4);
    out->n_dihedrals = (int32_t)(s->box.dih_p.size() / 3);
    out->n_impropers = (int32_t)(s->box.imp_p.size() / 2);
    out->n_pairs = (int32_t)(s->box.pair_p.size() / 2);
    return LIMA_OK;
}

lima_status lima_system_make_box(const lima_system* s, lima_box** out) {
    if (!s || !out) { g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return guarded([&] { *out = new lima_box(s->box); });
}

lima_status lima_system_layout(const lima_system* s, int32_t* gro_index) {
    if (!s || !gro_index) { g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    std::copy(s->layout_gro.begin(), s->layout_gro.end(), gro_index);
    return LIMA_OK;
}

lima_status lima_system_write_gro(const lima_system* s, const float* positions, const char* path) {
    if (!s || !positions || !path) { g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return guarded([&] {
        const std::string p(path);
        if (!ends_with(p, ".gro")) fail("expected .gro extension: " + p);
        const Gro& g = s->gro;
        std::vector<std::array<double, 3>> x(g.atoms.size());
        for (size_t i = 0; i < g.atoms.size(); i++)
            for (int d = 0; d < 3; d++) x[i][d] = g.atoms[i].x[d];
        const size_t nc32 = (size_t)s->box.nC * MAXP;
        size_t updated = 0;
        for (size_t e = 0; e < nc32; e++) {
            const int a = s->layout_gro[e];
            if (a < 0) continue;
            for (int d = 0; d < 3; d++) x[a][d] = positions[e * 3 + d];
            updated++;
        }
        for (size_t t = 0; t < s->tiny_atoms.size(); t++) {   // the molecule follows its first atom
            const auto& atoms = s->tiny_atoms[t];
            double delta[3];
            for (int d = 0; d < 3; d++) delta[d] = (double)positions[(nc32 + t) * 3 + d] - g.atoms[atoms[0]].x[d];
            for (int a : atoms) {
                for (int d = 0; d < 3; d++) x[a][d] = g.atoms[a].x[d] + delta[d];
                updated++;
            }
        }
        if (updated != g.atoms.size())
            fail("only " + std::to_string(updated) + " out of " + std::to_string(g.atoms.size()) + " particles were updated");
        std::ofstream o(p);
        if (!o) fail("failed to open " + p);
        o << g.title << "\n" << g.atoms.size() << "\n";
        for (size_t i = 0; i < g.atoms.size(); i++) write_gro_line(o, g.atoms[i], x[i].data());
        o.unsetf(std::ios::floatfield);
        o << std::setprecision(6) << (float)g.box[0] << " " << (float)g.box[1] << " " << (float)g.box[2] << "\n";
        if (!o) fail("write failed: " + p);
    });
}

}  // extern "C"
