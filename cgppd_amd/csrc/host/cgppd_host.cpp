// cgppd_host.cpp -- C++ host side (see cgppd_host.h).
#include "cgppd_host.h"

#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>

namespace cgppd {

// ---------------------------------------------------------------------------
// constants (float arithmetic as src/constants.h:8-26 evaluates it)
// ---------------------------------------------------------------------------
float rt_to_kcalmol()
{
    const float t_room = 298.0f, rgas = 8.3144621f, kcal = 4184.0f;
    return t_room * rgas / kcal;
}

float dh_conversion_factor()
{
    const float e = 1.602176487e-19f, n_a = 6.02214129e23f, eps0 = 8.85418782e-12f;
    const float kcal = 4184.0f, ang = 1e-10f, d_water = 80.0f;
    const float pi_f = (float)3.14159265358979323846264338327950288L;
    const float e2 = e * e;
    return e2 * n_a / (4.0f * pi_f * eps0 * kcal * ang * d_water);
}

float r0_constant() { return (float)std::pow(2.0, (double)(1.0f / 6.0f)); }

static std::runtime_error error(const std::string &what) { return std::runtime_error(what); }

// ---------------------------------------------------------------------------
// AminoAcids (src/AminoAcids.cpp:19-194)
// ---------------------------------------------------------------------------
void AminoAcids::init(const char *aminoAcidFile, const char *pairPotentialFile)
{
    clean = loadAminoAcidData(aminoAcidFile) && loadLJPotentialData(pairPotentialFile);
}

static bool getline_c(std::ifstream &in, std::string &line)
{
    if (!std::getline(in, line)) return false;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
}

bool AminoAcids::loadAminoAcidData(const char *filename)
{
    std::ifstream in(filename);
    if (!in.good()) throw error(std::string("Failed to open amino acid data file: ") + filename);
    std::string line;
    int i = 0;
    while (getline_c(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (i >= AA_COUNT) throw error("more than 20 amino acids in " + std::string(filename));
        AminoAcid &a = data[i];
        a.name = line.substr(0, 19);
        std::string s;
        if (!getline_c(in, s)) throw error("truncated amino acid entry");
        a.sname = s.substr(0, 3);
        if (!getline_c(in, s)) throw error("truncated amino acid entry");
        a.shortSymbol = s.empty() ? ' ' : s[0];
        if (!getline_c(in, s)) throw error("truncated amino acid entry");
        a.vanderWaalRadius = (float)std::atof(s.c_str());
        if (!getline_c(in, s)) throw error("truncated amino acid entry");
        float q = 0.f;
        std::sscanf(s.c_str(), "%f", &q);
        a.electrostaticCharge = q;
        a.index = i;
        ++i;
    }
    if (i != AA_COUNT) throw error("expected 20 amino acids in " + std::string(filename));
    return true;
}

bool AminoAcids::loadLJPotentialData(const char *filename)
{
    std::ifstream in(filename);
    if (!in.good()) throw error(std::string("Failed to open LJ potential data file: ") + filename);
    std::string line;
    int labels[AA_COUNT];
    bool labelline = true;
    int start = 0;
    while (getline_c(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string tok;
        if (labelline) {
            int n = 0;
            while (ss >> tok) {
                if (n >= AA_COUNT) throw error("too many labels in pair table");
                labels[n] = getAminoAcidIndex(tok.c_str());
                if (labels[n] < 0) throw error("unknown residue label in pair table: " + tok);
                ++n;
            }
            if (n != AA_COUNT) throw error("pair table label row needs 20 labels");
            labelline = false;
            continue;
        }
        if (start >= AA_COUNT) break;
        ss >> tok;  // row label
        for (int j = start; j < AA_COUNT; ++j) {
            if (!(ss >> tok)) throw error("pair table row too short");
            float e = 0.f;
            std::sscanf(tok.c_str(), "%f", &e);
            LJpotentials[labels[start]][labels[j]] = e;
            LJpotentials[labels[j]][labels[start]] = e;
        }
        ++start;
    }
    if (start != AA_COUNT) throw error("incomplete pair table " + std::string(filename));
    return true;
}

int AminoAcids::getAminoAcidIndex(const char *NAME) const
{
    const size_t n = std::strlen(NAME);
    if (n < 1) return -1;
    if (n == 3) {
        std::string key{(char)std::toupper(NAME[0]), (char)std::tolower(NAME[1]), (char)std::tolower(NAME[2])};
        for (int i = AA_COUNT - 1; i >= 0; --i)
            if (data[i].sname == key) return i;
        return -1;
    }
    if (n == 1) {
        const char c = (char)std::toupper(NAME[0]);
        for (int i = AA_COUNT - 1; i >= 0; --i)
            if (data[i].shortSymbol == c) return i;
        return -1;
    }
    std::string key(NAME);
    key[0] = (char)std::toupper(key[0]);
    for (int i = AA_COUNT - 1; i >= 0; --i)
        if (data[i].name == key) return i;
    return -1;
}

// ---------------------------------------------------------------------------
// Potential (src/Potential.cpp:5-42,116-230)
// ---------------------------------------------------------------------------
Potential::Potential(double LJ_kcal, double DH_kcal)
{
    LJ = LJ_kcal / rt_to_kcalmol();
    DH = DH_kcal / dh_conversion_factor();
}
Potential::Potential(double LJ_kcal, double DH_kcal, double bond_kcal, double angle_kcal, double torsion_kcal)
    : Potential(LJ_kcal, DH_kcal)
{
    bond = bond_kcal / (0.5 * 378);
    angle = std::exp(angle_kcal / -10.0);
    torsion = torsion_kcal;
}
double Potential::total_LJ() const { return LJ * rt_to_kcalmol(); }
double Potential::total_DH() const { return DH * dh_conversion_factor(); }
double Potential::total_bond() const { return bond * 0.5 * 378; }
double Potential::total_angle() const { return std::log(angle) * -10.0; }
double Potential::total_torsion() const { return torsion; }
double Potential::total() const
{
    // src/Potential.cpp:203-210
    return DH * dh_conversion_factor() + LJ * rt_to_kcalmol() + bond * 0.5 * 378 + std::log(angle) * -10.0 + torsion;
}
bool Potential::almost_equal(const Potential &p, float eps) const
{
    return std::fabs(LJ - p.LJ) < eps && std::fabs(DH - p.DH) < eps && std::fabs(bond - p.bond) < eps &&
           std::fabs(angle - p.angle) < eps && std::fabs(torsion - p.torsion) < eps;
}

bool TorsionalLookupMatrix::loadData(const char *filename, const AminoAcids &aminoAcids)
{
    std::ifstream in(filename);
    if (!in.good()) throw error(std::string("Failed to open torsional data file: ") + filename);
    std::string line;
    while (getline_c(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        char a0[8] = {0}, a1[8] = {0};
        double v = 0.0, sigma = 0.0;
        int n = 0;
        if (std::sscanf(line.c_str(), "%3s %3s %lf %d %lf", a0, a1, &v, &n, &sigma) != 5) continue;
        const int i = aminoAcids.getAminoAcidIndex(a0), j = aminoAcids.getAminoAcidIndex(a1);
        if (i < 0 || j < 0 || n < 1 || n > 4) throw error("bad torsion table line: " + line);
        V[i][j][n - 1] = v;
        Sigma[i][j][n - 1] = sigma / 57.29577951308232;  // degrees to radians
    }
    return true;
}

// Bonded terms with float vectors and double potentials (src/Potential.cpp:45-98).
static double bond_term(const Residue &a, const Residue &b, float L)
{
    const double r = b.distance(a, L);
    return (r - (double)3.81f) * (r - (double)3.81f);
}

static float fmagnitude(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

static double angle_term(const Residue &a, const Residue &b, const Residue &c)
{
    const float abx = a.position.x - b.position.x, aby = a.position.y - b.position.y, abz = a.position.z - b.position.z;
    const float cbx = c.position.x - b.position.x, cby = c.position.y - b.position.y, cbz = c.position.z - b.position.z;
    const double cmag = (double)(fmagnitude(cbx, cby, cbz) * fmagnitude(abx, aby, abz));
    float th = 0.0f;
    if (!(cmag < 1e-26)) th = (float)std::acos((double)(abx * cbx + aby * cby + abz * cbz) / cmag);
    const double t = th, da = t - (double)1.6f, db = t - (double)2.27f;
    return std::exp(-(double)0.1f * ((double)106.4f * da * da + (double)4.3f)) +
           std::exp(-(double)0.1f * ((double)26.3f * db * db));
}

static double torsion_term(const Residue &a, const Residue &b, const Residue &c, const Residue &d,
                           const TorsionalLookupMatrix &tor)
{
    const float b1x = b.position.x - a.position.x, b1y = b.position.y - a.position.y, b1z = b.position.z - a.position.z;
    const float b2x = c.position.x - b.position.x, b2y = c.position.y - b.position.y, b2z = c.position.z - b.position.z;
    const float b3x = d.position.x - c.position.x, b3y = d.position.y - c.position.y, b3z = d.position.z - c.position.z;
    const float x23 = b2y * b3z - b2z * b3y, y23 = b2z * b3x - b2x * b3z, z23 = b2x * b3y - b2y * b3x;
    const float x12 = b1y * b2z - b1z * b2y, y12 = b1z * b2x - b1x * b2z, z12 = b1x * b2y - b1y * b2x;
    const double phi = (double)std::atan2(fmagnitude(b2x, b2y, b2z) * (b1x * x23 + b1y * y23 + b1z * z23),
                                          x12 * x23 + y12 * y23 + z12 * z23);
    const int j = b.aminoAcidIndex, k = c.aminoAcidIndex;
    double e = 0.0;
    for (int n = 1; n <= 4; ++n) e += (1 + std::cos(n * phi - tor.getSigma(j, k, n))) * tor.getV(j, k, n);
    return e;
}

double calculate_LJ(int ti, int tj, double r, const AminoAcids &aa)
{
    const double Eij = (double)(0.159f * (aa.LJpotentials[ti][tj] - (-2.27f)));
    const double sij = (double)(0.5f * (aa.data[ti].vanderWaalRadius + aa.data[tj].vanderWaalRadius));
    const double t = std::pow(sij / r, 6.0);
    double lj = -4.0 * Eij * t * (t - 1.0);
    if (Eij > 0.0 && r < sij * (double)r0_constant()) lj = -lj + 2.0 * Eij;
    return lj;
}

double calculate_DH(float qi, float qj, double r)
{
    const float e = std::exp((float)(-r / 10.0));
    return (double)(qi * qj * e) / r;
}

double Residue::distance(const Residue &o, float L) const
{
    float X = position.x - o.position.x, Y = position.y - o.position.y, Z = position.z - o.position.z;
    if (L != 0.0f) {
        X = X - L * std::round(X / L);
        Y = Y - L * std::round(Y / L);
        Z = Z - L * std::round(Z / L);
    }
    return (double)std::sqrt(X * X + Y * Y + Z * Z);
}

// ---------------------------------------------------------------------------
// Molecule (src/Molecule.cpp:25-64,146-208,506-612)
// ---------------------------------------------------------------------------
static float scan_float(const std::string &line, size_t col)
{
    float v = 0.f;
    if (col < line.size()) std::sscanf(line.c_str() + col, "%f", &v);
    return v;
}

void Molecule::initFromPDB(const char *pdbfilename, const AminoAcids &aa)
{
    filename = pdbfilename;
    std::ifstream in(pdbfilename);
    if (!in.good()) throw error(std::string("Failed to open PDB file: ") + pdbfilename);
    Residues.clear();
    chainCount = 0;
    std::string line;
    while (getline_c(in, line)) {
        if (line.compare(0, 4, "ATOM") == 0) {
            char name[16] = {0}, resName[16] = {0};
            if (line.size() > 12) std::sscanf(line.c_str() + 12, "%15s", name);
            if (std::strcmp(name, "CA") != 0) continue;
            if (line.size() > 17) std::sscanf(line.c_str() + 17, "%15s", resName);
            Residue R;
            R.aminoAcidIndex = aa.getAminoAcidIndex(resName);
            if (R.aminoAcidIndex < 0) throw error(std::string("unknown residue ") + resName + " in " + pdbfilename);
            R.electrostaticCharge = aa.get(R.aminoAcidIndex).electrostaticCharge;
            R.vanderWaalRadius = aa.get(R.aminoAcidIndex).vanderWaalRadius;
            R.chainId = line.size() > 21 ? line[21] : ' ';
            R.chain_int_id = chainCount;
            if (line.size() > 22) std::sscanf(line.c_str() + 22, "%d", &R.resSeq);
            R.position.x = scan_float(line, 30);
            R.position.y = scan_float(line, 38);
            R.position.z = scan_float(line, 46);
            R.moleculeId = index;
            Residues.push_back(R);
        } else if (line.compare(0, 3, "TER") == 0) {
            ++chainCount;
        } else if (line.compare(0, 3, "END") == 0) {
            break;
        }
    }
    if (!Residues.empty() && chainCount == 0) chainCount = 1;
    residueCount = (int)Residues.size();
    if (residueCount == 0) throw error(std::string("no C-alpha atoms in ") + pdbfilename);
    // src/Molecule.cpp:590-609: float accumulation of the centre
    Vector3f c;
    for (const Residue &r : Residues) {
        c.x += r.position.x;
        c.y += r.position.y;
        c.z += r.position.z;
    }
    c.x /= residueCount;
    c.y /= residueCount;
    c.z /= residueCount;
    center = c;
    for (Residue &r : Residues) {
        r.relativePosition.x = r.position.x - c.x;
        r.relativePosition.y = r.position.y - c.y;
        r.relativePosition.z = r.position.z - c.z;
    }
}

void Molecule::setPosition(Vector3f c)
{
    center = c;
    for (Residue &r : Residues) {
        r.position.x = r.relativePosition.x + c.x;
        r.position.y = r.relativePosition.y + c.y;
        r.position.z = r.relativePosition.z + c.z;
    }
}

void Molecule::translate(Vector3f v)
{
    Vector3f c{center.x + v.x, center.y + v.y, center.z + v.z};
    if (bounding_value != 0.0f) {
        c.x = std::fmod(c.x, bounding_value);
        c.y = std::fmod(c.y, bounding_value);
        c.z = std::fmod(c.z, bounding_value);
    }
    setPosition(c);
}

void Molecule::rotate(const double axis[3], double angle)
{
    // src/Quaternion.cpp:12-38,95-112
    double w = std::cos(angle / 2.0), s = std::sin(angle / 2.0);
    double x = s * axis[0], y = s * axis[1], z = s * axis[2];
    const double m = std::sqrt(w * w + x * x + y * y + z * z);
    w /= m;
    x /= m;
    y /= m;
    z /= m;
    const double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    const double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    for (Residue &r : Residues) {
        const double vx = r.relativePosition.x, vy = r.relativePosition.y, vz = r.relativePosition.z;
        r.relativePosition.x = (float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0);
        r.relativePosition.y = (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2));
        r.relativePosition.z = (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2));
        r.position.x = center.x + r.relativePosition.x;
        r.position.y = center.y + r.relativePosition.y;
        r.position.z = center.z + r.relativePosition.z;
    }
}

static std::vector<int> topology_list(const std::vector<int> &code, const std::vector<int> &chain, bool all_flexible,
                                      const std::vector<std::vector<int>> &segments, int kind)
{
    std::vector<int> off{0}, res;
    for (const auto &sg : segments) {
        res.insert(res.end(), sg.begin(), sg.end());
        off.push_back((int)res.size());
    }
    if (res.empty()) res.push_back(0);
    int n = 0;
    int rc = cgppd_topology_list((int)code.size(), code.data(), chain.data(), all_flexible ? 1 : 0,
                                 (int)segments.size(), off.data(), res.data(), kind, nullptr, &n);
    if (rc) throw error("cgppd_topology_list failed (bad segment?)");
    std::vector<int> out(std::max(n, 1));
    cgppd_topology_list((int)code.size(), code.data(), chain.data(), all_flexible ? 1 : 0, (int)segments.size(),
                        off.data(), res.data(), kind, out.data(), &n);
    out.resize(n);
    return out;
}

void Molecule::init_graph()
{
    // src/Molecule.cpp:57-61 (Graph::init); the library's builder restates
    // src/Geometry.cpp:34-192
    is_flexible = all_flexible || !segments.empty();
    bonds.clear();
    angles.clear();
    torsions.clear();
    indirect.clear();
    domain_of.assign(residueCount, -1);
    segbond_of.assign(residueCount, -1);
    std::vector<int> code, chain;
    for (const Residue &r : Residues) {
        code.push_back((unsigned char)r.chainId);
        chain.push_back(r.chain_int_id);
    }
    bonds = topology_list(code, chain, all_flexible, segments, 0);
    angles = topology_list(code, chain, all_flexible, segments, 1);
    torsions = topology_list(code, chain, all_flexible, segments, 2);
    indirect = topology_list(code, chain, all_flexible, segments, 7);
    domain_of = topology_list(code, chain, all_flexible, segments, 8);
    segbond_of = topology_list(code, chain, all_flexible, segments, 13);
    vertex_count = topology_list(code, chain, all_flexible, segments, 14)[0];
}

Potential Molecule::E(bool include_LJ_and_DH, const AminoAcids &aa, const TorsionalLookupMatrix *tor) const
{
    if (!tor) tor = torsion_data;
    Potential p;
    if (!is_flexible) return p;
    const int n = residueCount;
    auto indirect_pair = [&](int i, int j) {
        for (size_t k = 0; k + 1 < indirect.size(); k += 2)
            if (indirect[k] == i && indirect[k + 1] == j) return true;
        return false;
    };
    if (include_LJ_and_DH) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                const Residue &a = Residues[i], &b = Residues[j];
                if (a.same_rigid_domain_as(b) || a.chain_neighbour_of(b) || a.bonded_to(b) || indirect_pair(i, j))
                    continue;
                const double r = a.distance(b, bounding_value) + (double)EPS;
                p.increment_LJ(calculate_LJ(a.aminoAcidIndex, b.aminoAcidIndex, r, aa));
                p.increment_DH(calculate_DH(a.electrostaticCharge, b.electrostaticCharge, r));
            }
    } else {
        for (size_t k = 0; k + 1 < indirect.size(); k += 2) {
            const Residue &a = Residues[indirect[k]], &b = Residues[indirect[k + 1]];
            const double r = a.distance(b, bounding_value) + (double)EPS;
            p.increment_LJ(-calculate_LJ(a.aminoAcidIndex, b.aminoAcidIndex, r, aa));
            p.increment_DH(-calculate_DH(a.electrostaticCharge, b.electrostaticCharge, r));
        }
    }
    for (size_t k = 0; k + 1 < bonds.size(); k += 2)
        p.increment_bond(bond_term(Residues[bonds[k]], Residues[bonds[k + 1]], bounding_value));
    for (size_t k = 0; k + 2 < angles.size(); k += 3)
        p.increment_angle(angle_term(Residues[angles[k]], Residues[angles[k + 1]], Residues[angles[k + 2]]));
    if (!tor && !torsions.empty()) throw error("flexible molecule without a torsion table");
    for (size_t k = 0; k + 3 < torsions.size(); k += 4)
        p.increment_torsion(torsion_term(Residues[torsions[k]], Residues[torsions[k + 1]], Residues[torsions[k + 2]],
                                         Residues[torsions[k + 3]], *tor));
    return p;
}

void Molecule::init(const moldata &mol, const AminoAcids &aa, int idx, float bv, const TorsionalLookupMatrix *tor)
{
    index = idx;
    bounding_value = bv;
    name = mol.name;
    torsion_data = tor;
    all_flexible = mol.all_flexible;
    segments = mol.segments;
    initFromPDB(mol.pdbfilename.c_str(), aa);
    if (mol.translate) {
        if (mol.px || mol.py || mol.pz) translate(Vector3f{mol.px, mol.py, mol.pz});
    } else {
        setPosition(Vector3f{mol.px, mol.py, mol.pz});
    }
    if (mol.ra) {
        double v[3] = {mol.rx, mol.ry, mol.rz};
        const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (n > 0) {
            v[0] /= n;
            v[1] /= n;
            v[2] /= n;
            rotate(v, mol.ra);
        }
    }
    crowder = mol.crowder;
    init_graph();
}

// ---------------------------------------------------------------------------
// DeviceContext (replaces setup_CUDA / teardown_CUDA)
// ---------------------------------------------------------------------------
DeviceContext::DeviceContext(int device, float box_dimension, const AminoAcids &aa)
{
    int rc = cgppd_create(&ctx_, device);
    if (rc) throw error("cgppd_create failed on device " + std::to_string(device) + " (status " + std::to_string(rc) + ")");
    float eps[AA_COUNT * AA_COUNT], sigma[AA_COUNT], charge[AA_COUNT];
    for (int i = 0; i < AA_COUNT; ++i) {
        sigma[i] = aa.data[i].vanderWaalRadius;
        charge[i] = aa.data[i].electrostaticCharge;
        for (int j = 0; j < AA_COUNT; ++j) eps[i * AA_COUNT + j] = aa.LJpotentials[i][j];
    }
    check(cgppd_set_force_field(ctx_, eps, sigma, charge), "cgppd_set_force_field");
    check(cgppd_set_box(ctx_, box_dimension), "cgppd_set_box");
}

DeviceContext::~DeviceContext()
{
    if (ctx_) cgppd_destroy(ctx_);
}

void DeviceContext::check(int rc, const char *what) const
{
    if (rc != CGPPD_OK) throw error(std::string(what) + ": " + cgppd_last_error(ctx_));
}

// ---------------------------------------------------------------------------
// Replica
// ---------------------------------------------------------------------------
void Replica::loadMolecule(const moldata &mol)
{
    Molecule m;
    m.init(mol, aminoAcids, moleculeCount, boundingValue, &torsions);
    residueCount += m.residueCount;
    max_residue_count = std::max(max_residue_count, m.residueCount);
    if (!mol.crowder) {
        ++nonCrowderCount;
        nonCrowderResidues += m.residueCount;
    }
    molecules.push_back(std::move(m));
    ++moleculeCount;
}

void Replica::init_first_replica(const argdata &parameters, const AminoAcids &amino_acid_data, int)
{
    aminoAcids = amino_acid_data;
    boundingValue = parameters.bound;
    repulsiveCrowding = parameters.repulsiveCrowding;
    molecules.clear();
    moleculeCount = residueCount = nonCrowderCount = nonCrowderResidues = max_residue_count = 0;
    bool flexible = false;
    for (const moldata &m : parameters.mdata) flexible = flexible || m.all_flexible || !m.segments.empty();
    if (flexible) torsions.loadData(parameters.torsionFile.c_str(), amino_acid_data);
    for (const moldata &m : parameters.mdata) loadMolecule(m);
    // replica-wide UIDs from 1 (src/Replica.cpp:30-38, src/Geometry.cpp:256-302)
    int chain_offset = 1, domain_offset = 1, bond_offset = 1;
    for (Molecule &mol : molecules) {
        int n_domains = 0, n_bonds = 0;
        for (int d : mol.domain_of) n_domains = std::max(n_domains, d + 1);
        for (int b : mol.segbond_of) n_bonds = std::max(n_bonds, b + 1);
        for (int i = 0; i < mol.residueCount; ++i) {
            Residue &r = mol.Residues[i];
            r.rigid_domain_UID = mol.domain_of[i] >= 0 ? domain_offset + mol.domain_of[i] : 0;
            r.segment_bond_UID = mol.segbond_of[i] >= 0 ? bond_offset + mol.segbond_of[i] : 0;
            r.chain_UID = i < mol.vertex_count ? chain_offset + r.chain_int_id : 0;
        }
        chain_offset += mol.chainCount;
        domain_offset += n_domains;
        bond_offset += n_bonds;
    }
    // the device requires non-crowders first (src/Replica.cpp:924)
    bool seen_crowder = false;
    for (const Molecule &m : molecules) {
        if (m.crowder) seen_crowder = true;
        else if (seen_crowder) throw error("non-crowder molecules must precede crowders");
    }
}

Potential Replica::internal_molecule_E(bool include_LJ_and_DH) const
{
    // src/Replica.cpp:532-560
    Potential p;
    for (const Molecule &m : molecules)
        if (!m.crowder) p += m.E(include_LJ_and_DH, aminoAcids, &torsions);
    return p;
}

bool Replica::has_flexible() const
{
    for (const Molecule &m : molecules)
        if (m.is_flexible) return true;
    return false;
}

void Replica::copy(const Replica &r)
{
    aminoAcids = r.aminoAcids;
    torsions = r.torsions;
    molecules = r.molecules;
    for (Molecule &m : molecules) m.torsion_data = &torsions;
    moleculeCount = r.moleculeCount;
    residueCount = r.residueCount;
    nonCrowderCount = r.nonCrowderCount;
    nonCrowderResidues = r.nonCrowderResidues;
    max_residue_count = r.max_residue_count;
    boundingValue = r.boundingValue;
    label = r.label;
    temperature = r.temperature;
    potential = r.potential;
    repulsiveCrowding = r.repulsiveCrowding;
    calculate_rigid_potential_only = r.calculate_rigid_potential_only;
}

void Replica::init_child_replica(const Replica &ir, int index, double geometricTemperature, double geometricRotation,
                                 double geometricTranslate, const argdata &parameters)
{
    copy(ir);
    label = (short)index;
    temperature = (float)(parameters.temperatureMin * std::pow(geometricTemperature, index));
    rotateStep = MIN_ROTATION * std::pow(geometricRotation, index);
    translateStep = (float)(MIN_TRANSLATION * std::pow(geometricTranslate, index));
}

void Replica::exchangeReplicas(Replica &r)
{
    std::swap(label, r.label);
    std::swap(temperature, r.temperature);
    std::swap(accept, r.accept);
    std::swap(acceptA, r.acceptA);
    std::swap(reject, r.reject);
    std::swap(totalAcceptReject, r.totalAcceptReject);
    std::swap(totalAccept, r.totalAccept);
    std::swap(acceptanceRatio, r.acceptanceRatio);
    std::swap(accumulativeAcceptanceRatio, r.accumulativeAcceptanceRatio);
    std::swap(boundSamples, r.boundSamples);
    std::swap(samplesSinceLastExchange, r.samplesSinceLastExchange);
    std::swap(totalBoundSamples, r.totalBoundSamples);
    std::swap(totalSamples, r.totalSamples);
    std::swap(fractionBound, r.fractionBound);
    std::swap(accumulativeFractionBound, r.accumulativeFractionBound);
}

Potential Replica::E(const Molecule *a, const Molecule *b) const
{
    Potential p;
    for (const Residue &ri : a->Residues)
        for (const Residue &rj : b->Residues) {
            const double r = ri.distance(rj, boundingValue) + (double)EPS;
            p.increment_LJ(calculate_LJ(ri.aminoAcidIndex, rj.aminoAcidIndex, r, aminoAcids));
            p.increment_DH(calculate_DH(ri.electrostaticCharge, rj.electrostaticCharge, r));
        }
    return p;
}

Potential Replica::E() const
{
    // src/Replica.cpp:432-493 (one running Potential over all pairs); under
    // repulsive crowding a pair with a crowder adds (6/r)^12 [RT] below the
    // 15 A cutoff and no DH (src/Replica.cpp:416-419,436-470,
    // src/definitions.h:38)
    Potential p;
    for (int I = 0; I < moleculeCount; ++I)
        for (int J = I + 1; J < moleculeCount; ++J) {
            const bool crowd = repulsiveCrowding && (molecules[I].crowder || molecules[J].crowder);
            for (const Residue &ri : molecules[I].Residues)
                for (const Residue &rj : molecules[J].Residues) {
                    const double r = ri.distance(rj, boundingValue) + (double)EPS;
                    if (crowd) {
                        if (r < 15.0f) p.increment_LJ(powf(6.0f / (float)r, 12.0f));
                        continue;
                    }
                    p.increment_LJ(calculate_LJ(ri.aminoAcidIndex, rj.aminoAcidIndex, r, aminoAcids));
                    p.increment_DH(calculate_DH(ri.electrostaticCharge, rj.electrostaticCharge, r));
                }
        }
    return p;
}

std::vector<int> Replica::molecule_offsets() const
{
    std::vector<int> off(moleculeCount + 1, 0);
    for (int m = 0; m < moleculeCount; ++m) off[m + 1] = off[m] + molecules[m].residueCount;
    return off;
}

std::vector<int> Replica::residue_types() const
{
    std::vector<int> t;
    for (const Molecule &m : molecules)
        for (const Residue &r : m.Residues) t.push_back(r.aminoAcidIndex);
    return t;
}

std::vector<float> Replica::positions() const
{
    std::vector<float> x;
    for (const Molecule &m : molecules)
        for (const Residue &r : m.Residues) x.insert(x.end(), {r.position.x, r.position.y, r.position.z});
    return x;
}

std::vector<float> Replica::relative_positions() const
{
    std::vector<float> x;
    for (const Molecule &m : molecules)
        for (const Residue &r : m.Residues)
            x.insert(x.end(), {r.relativePosition.x, r.relativePosition.y, r.relativePosition.z});
    return x;
}

std::vector<float> Replica::centres() const
{
    std::vector<float> c;
    for (const Molecule &m : molecules) c.insert(c.end(), {m.center.x, m.center.y, m.center.z});
    return c;
}

void Replica::setup_device(DeviceContext *dev, int slot)
{
    dev_ = dev;
    slot_ = slot;
}

void Replica::ReplicaDataToDevice()
{
    if (!dev_) throw error("Replica::ReplicaDataToDevice: no device context (call setup_device)");
    // src/Replica.cpp:753-757: the rigid-only override travels with the data
    // (the device context holds one flag for all of its replicas)
    dev_->check(cgppd_set_rigid_potential_only(dev_->ctx(), calculate_rigid_potential_only ? 1 : 0),
                "cgppd_set_rigid_potential_only");
    const std::vector<float> x = positions();
    dev_->check(cgppd_upload_positions(dev_->ctx(), slot_, x.data()), "cgppd_upload_positions");
}

void Replica::MoleculeDataToDevice(int moleculeId)
{
    if (!dev_) throw error("Replica::MoleculeDataToDevice: no device context");
    std::vector<float> x;
    for (const Residue &r : molecules.at(moleculeId).Residues)
        x.insert(x.end(), {r.position.x, r.position.y, r.position.z});
    dev_->check(cgppd_update_molecule(dev_->ctx(), slot_, moleculeId, x.data()), "cgppd_update_molecule");
}

double Replica::EonDevice()
{
    if (!dev_) throw error("Replica::EonDevice: no device context");
    double t = 0.0;
    dev_->check(cgppd_energy(dev_->ctx(), slot_, nullptr, nullptr, &t), "cgppd_energy");
    return t;
}

double Replica::EonDeviceNC()
{
    if (!dev_) throw error("Replica::EonDeviceNC: no device context");
    double t = 0.0;
    dev_->check(cgppd_energy_nc(dev_->ctx(), slot_, nullptr, nullptr, &t), "cgppd_energy_nc");
    return t;
}

void Replica::EonDeviceAsync()
{
    if (!dev_) throw error("Replica::EonDeviceAsync: no device context");
    dev_->check(cgppd_energy_async(dev_->ctx(), slot_), "cgppd_energy_async");
}

double Replica::SumGridResults()
{
    if (!dev_) throw error("Replica::SumGridResults: no device context");
    double t = 0.0;
    dev_->check(cgppd_energy_wait(dev_->ctx(), slot_, nullptr, nullptr, &t), "cgppd_energy_wait");
    return t;
}

void Replica::fraction_bound(FILE *f)
{
    // src/Replica.cpp:1095-1107
    totalBoundSamples += boundSamples;
    totalSamples += samplesSinceLastExchange;
    fractionBound = float(boundSamples) / std::max(1.0f, float(samplesSinceLastExchange));
    accumulativeFractionBound = float(totalBoundSamples) / std::max(1.0f, float(totalSamples));
    if (f) std::fprintf(f, "| %6.4f %6.4f ", fractionBound, accumulativeFractionBound);
    boundSamples = 0;
    samplesSinceLastExchange = 0;
}

void Replica::acceptance_ratio(FILE *f)
{
    // src/Replica.cpp:1109-1122 (no guard against an empty interval: 0/0
    // prints as nan there too)
    totalAccept += acceptA + accept;
    totalAcceptReject += acceptA + accept + reject;
    acceptanceRatio = float(acceptA + accept) / float(acceptA + accept + reject);
    accumulativeAcceptanceRatio = float(totalAccept) / float(totalAcceptReject);
    if (f) std::fprintf(f, "| %6.4f %6.4f ", acceptanceRatio, accumulativeAcceptanceRatio);
    acceptA = accept = reject = 0;
}

std::string Replica::sample(const char *prefix, int current_step, float boundEnergyThreshHold)
{
    // src/Replica.cpp:1037-1093; the reference records every sample
    // (bool_recordAllSamples, src/definitions.h:36)
    ++samplesSinceLastExchange;
    if (nonCrowderPotential < boundEnergyThreshHold) ++boundSamples;
    char savename[512];
    std::snprintf(savename, sizeof(savename), "output/%s/pdb/sample_%d_%0.1fK_%5.2f.pdb", prefix, current_step,
                  temperature, nonCrowderPotential);
    saveAsSinglePDB(savename, prefix, current_step);
    char record[1024];
    std::snprintf(record, sizeof(record), "%d; %0.5f (%0.5f); %0.1f\n%s\n", current_step, nonCrowderPotential,
                  potential, temperature, savename);
    return record;
}

void Replica::saveAsSinglePDB(const char *filename, const char *prefix, int sample, bool skip_crowders) const
{
    // src/Replica.cpp:1124-1168: chain identifiers from the residues' chain
    // UIDs ('@' + UID), a TER record where the chain changes and after
    // every molecule
    FILE *out = std::fopen(filename, "w");
    if (!out) throw error(std::string("cannot write ") + filename);
    std::fprintf(out, "REMARK simulation: %s \n", prefix);
    std::fprintf(out, "REMARK temperature: %5.1f \n", temperature);
    std::fprintf(out, "REMARK sample: %d \n", sample);
    std::fprintf(out, "REMARK potential: %0.10f \n", float(potential));
    int item = 0, last_seq = 0;
    char chain = 0, last_chain = 0;
    for (int m = 0; m < moleculeCount; ++m) {
        if (skip_crowders && m >= nonCrowderCount) break;
        const Molecule &mol = molecules[m];
        std::string last_name;
        for (size_t i = 0; i < mol.Residues.size(); ++i) {
            const Residue &r = mol.Residues[i];
            chain = (char)(64 + r.chain_UID);
            if (i > 0 && chain != last_chain)
                std::fprintf(out, "TER   %5d      %3s %c%4d \n", item, last_name.c_str(), last_chain, last_seq);
            ++item;
            last_name = aminoAcids.get(r.aminoAcidIndex).getSNAME();
            std::fprintf(out, "ATOM  %5d CA   %3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n", item, last_name.c_str(),
                         chain, r.resSeq, r.position.x, r.position.y, r.position.z, 0.0f, 0.0f);
            last_seq = r.resSeq;
            last_chain = chain;
        }
        std::fprintf(out, "TER   %5d      %3s %c%4d \n", item, last_name.c_str(), chain, last_seq);
    }
    std::fprintf(out, "END \n");
    std::fclose(out);
}

// ---------------------------------------------------------------------------
// Simulation
// ---------------------------------------------------------------------------
Simulation::~Simulation() { closeSamplingFiles(); }

void Simulation::printHelp() const
{
    std::printf(
        "Usage: cgppd -f <config file> [options]\n"
        "\t-h|--help\n\t-f|--file <config>\n\t-v|--verbosity <0-4>\n\t-q|--nosim\n"
        "\t-t|--threads n (accepted, unused: replicas run on the device)\n"
        "\t-s|--streams n (accepted, unused)\n\t-g|--gpus n (one GPU)\n\t-z|--gpuoffset n (device id)\n"
        "\t-m|--mcsteps n\n\t-a|--sampleafter n\n\t-e|--resteps n\n\t-r|--replicas n\n\t-o|--output prefix\n"
        "\t-b|--boundary L\n\t-x|--tmax T\n\t-n|--tmin T\n\t-d|--blockdim n (accepted, unused)\n"
        "\t-S|--seed n (device Philox seed)\n\t-G|--nograph (accepted, unused)\n");
}

void Simulation::getArgs(int argc, char **argv, bool first_pass)
{
    // src/Simulation.cpp:78-263: two passes, the file first
    static const struct option long_options[] = {
        {"file", required_argument, 0, 'f'},      {"help", no_argument, 0, 'h'},
        {"verbosity", required_argument, 0, 'v'}, {"preview", no_argument, 0, 'p'},
        {"nosim", no_argument, 0, 'q'},           {"threads", required_argument, 0, 't'},
        {"streams", required_argument, 0, 's'},   {"gpus", required_argument, 0, 'g'},
        {"gpuoffset", required_argument, 0, 'z'}, {"mcsteps", required_argument, 0, 'm'},
        {"sampleafter", required_argument, 0, 'a'}, {"resteps", required_argument, 0, 'e'},
        {"replicas", required_argument, 0, 'r'},  {"output", required_argument, 0, 'o'},
        {"boundary", required_argument, 0, 'b'},  {"tmax", required_argument, 0, 'x'},
        {"tmin", required_argument, 0, 'n'},      {"blockdim", required_argument, 0, 'd'},
        {"seed", required_argument, 0, 'S'},      {"nograph", no_argument, 0, 'G'},
        {0, 0, 0, 0},
    };
    const char short_options[] = "hf:pqt:s:g:z:m:a:e:r:o:b:x:n:d:v:w:S:G";
    int opt_index = 0;
    optind = 1;
    if (first_pass) {
        while (true) {
            const int opt = getopt_long(argc, argv, short_options, long_options, &opt_index);
            if (opt == -1) break;
            if (opt == 'h') {
                printHelp();
                throw error("help");
            }
            if (opt == 'f') {
                parameters.file = optarg;
                parameters.inputFile = true;
            }
        }
        if (!parameters.inputFile) {
            printHelp();
            throw error("No configuration file provided.");
        }
        return;
    }
    while (true) {
        const int opt = getopt_long(argc, argv, short_options, long_options, &opt_index);
        if (opt == -1) break;
        switch (opt) {
        case 'q': parameters.skipsimulation = true; break;
        case 't': parameters.threads = std::atoi(optarg); break;
        case 's': parameters.streams = std::atoi(optarg); break;
        case 'g': parameters.gpus = std::atoi(optarg); break;
        case 'z': parameters.gpuoffset = std::atoi(optarg); break;
        case 'm': parameters.MCsteps = std::atoi(optarg); break;
        case 'a': parameters.sampleStartsAfter = std::atoi(optarg); break;
        case 'e': parameters.REsteps = std::atoi(optarg); break;
        case 'r': parameters.replicas = std::atoi(optarg); break;
        case 'o': parameters.prefix = std::string(optarg) + "_" + std::to_string(parameters.pid); break;
        case 'b': parameters.bound = (float)std::atof(optarg); break;
        case 'x': parameters.temperatureMax = (float)std::atof(optarg); break;
        case 'n': parameters.temperatureMin = (float)std::atof(optarg); break;
        case 'd':
            parameters.cuda_blockSize = std::atoi(optarg);
            parameters.auto_blockdim = false;
            break;
        case 'S': parameters.seed = std::strtoull(optarg, nullptr, 10); break;
        case 'G': parameters.use_graph = false; break;
        default: break;
        }
    }
}

void Simulation::loadArgsFromFile()
{
    // src/Simulation.cpp:265-453
    std::ifstream in(parameters.file);
    if (!in.good()) throw error("Failed to open config file: " + parameters.file);
    enum { PARAMS, MOLECULES, CROWDERS, SEGMENTS } section = PARAMS;
    std::string line;
    while (getline_c(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (line == "files") {
            section = MOLECULES;
        } else if (line == "crowders") {
            section = CROWDERS;
        } else if (line == "segments") {
            section = SEGMENTS;
        } else if (section == MOLECULES || section == CROWDERS) {
            moldata m;
            char flag = 0, pdb[512] = {0}, name[512] = {0};
            const int n = std::sscanf(line.c_str(), "%c(%f,%f,%f) r(%f,%f,%f,%f) %511s %511s", &flag, &m.px, &m.py,
                                      &m.pz, &m.rx, &m.ry, &m.rz, &m.ra, pdb, name);
            if (n < 8) {
                std::fprintf(stderr, "warning: unable to parse molecule: %s\n", line.c_str());
                continue;
            }
            m.translate = flag == 't';
            m.pdbfilename = pdb;
            m.name = name;
            m.crowder = section == CROWDERS;
            if (section == MOLECULES) ++parameters.nonCrowders;
            parameters.mdata.push_back(m);
            if (!m.name.empty()) parameters.mdata_map[m.name] = (int)parameters.mdata.size() - 1;
        } else if (section == SEGMENTS) {
            std::istringstream ss(line);
            std::string molname, word;
            ss >> molname >> word;
            auto it = parameters.mdata_map.find(molname);
            if (it == parameters.mdata_map.end()) throw error("segment for unknown molecule " + molname);
            if (word == "all") {
                parameters.mdata[it->second].all_flexible = true;
            } else {
                std::vector<int> seg;
                do {
                    seg.push_back(std::atoi(word.c_str()) - 1);
                } while (ss >> word);
                parameters.mdata[it->second].segments.push_back(seg);
            }
        } else {
            std::istringstream ss(line);
            std::string key, value;
            ss >> key >> value;
            if (key == "gpus") parameters.gpus = std::atoi(value.c_str());
            else if (key == "gpuoffset") parameters.gpuoffset = std::atoi(value.c_str());
            else if (key == "streams") parameters.streams = std::atoi(value.c_str());
            else if (key == "threads") parameters.threads = std::atoi(value.c_str());
            else if (key == "mcsteps") parameters.MCsteps = std::atoi(value.c_str());
            else if (key == "resteps") parameters.REsteps = std::atoi(value.c_str());
            else if (key == "temperaturemax") parameters.temperatureMax = (float)std::atof(value.c_str());
            else if (key == "temperaturemin") parameters.temperatureMin = (float)std::atof(value.c_str());
            else if (key == "boundary") parameters.bound = (float)std::atof(value.c_str());
            else if (key == "replicas") parameters.replicas = std::atoi(value.c_str());
            else if (key == "samplefrequency") parameters.sampleFrequency = std::atoi(value.c_str());
            else if (key == "sampleafter") parameters.sampleStartsAfter = std::atoi(value.c_str());
            else if (key == "prefix") parameters.prefix = value + "_" + std::to_string(parameters.pid);
            else if (key == "title") {
                std::string rest, w;
                rest = value;
                while (ss >> w) rest += " " + w;
                parameters.title = rest;
            }
            else if (key == "seed") parameters.seed = std::strtoull(value.c_str(), nullptr, 10);
            else if (key == "repulsivecrowding") parameters.repulsiveCrowding = std::atoi(value.c_str()) != 0;
            else if (key == "aminoacids") parameters.aminoAcidFile = value;
            else if (key == "pairpotentials") parameters.pairPotentialFile = value;
            else if (key == "torsions") parameters.torsionFile = value;
            else std::fprintf(stderr, "warning: unknown parameter: %s\n", key.c_str());
        }
    }
}

void Simulation::check_and_modify_parameters()
{
    // src/Simulation.cpp:455-514 (thread/stream checks no longer apply)
    if (parameters.bound <= 0) parameters.bound = 150.0f;
    if (parameters.temperatureMax < parameters.temperatureMin)
        std::swap(parameters.temperatureMax, parameters.temperatureMin);
    if (parameters.replicas < 1) throw error("need at least one replica");
    if (parameters.REsteps < 1) parameters.REsteps = 1;
    if (parameters.gpus != 1) {
        std::fprintf(stderr, "warning: this build runs on one GPU; using device %d\n", parameters.gpuoffset);
        parameters.gpus = 1;
    }
}

void Simulation::init(int argc, char **argv, int pid)
{
    parameters.pid = pid;
    getArgs(argc, argv, true);
    loadArgsFromFile();
    getArgs(argc, argv, false);
    check_and_modify_parameters();
    aminoAcidData.init(parameters.aminoAcidFile.c_str(), parameters.pairPotentialFile.c_str());
    initialReplica.init_first_replica(parameters, aminoAcidData, 30);
    if (initialReplica.moleculeCount == 0) throw error("No molecules loaded.");
    initialReplica.potential = initialReplica.E().total();
    const int R = parameters.replicas;
    // src/Simulation.cpp:590-592
    geometricTemperature = std::pow(double(parameters.temperatureMax / parameters.temperatureMin), 1.0 / double(std::max(R - 1, 1)));
    geometricTranslate = std::pow(double(MAX_TRANSLATION / MIN_TRANSLATION), 1.0 / double(std::max(R - 1, 1)));
    geometricRotation = std::pow(double(MAX_ROTATION / MIN_ROTATION), 1.0 / double(std::max(R - 1, 1)));
    replica.assign(R, Replica());
    for (int i = 0; i < R; ++i)
        replica[i].init_child_replica(initialReplica, i, geometricTemperature, geometricRotation, geometricTranslate,
                                      parameters);
    position_of_temperature.clear();
    temperature.clear();
    for (int i = 0; i < R; ++i) {
        position_of_temperature[replica[i].temperature] = i;
        temperature[i] = replica[i].temperature;
    }
    near_300k = temperature[0];
    for (auto &p : position_of_temperature)
        if (std::fabs(p.first - 300.0f) < std::fabs(near_300k - 300.0f)) near_300k = p.first;
}

void Simulation::setup_device(int device)
{
    dev_.reset(new DeviceContext(device, parameters.bound, aminoAcidData));
    const std::vector<int> off = initialReplica.molecule_offsets();
    const std::vector<int> types = initialReplica.residue_types();
    dev_->check(cgppd_set_system(dev_->ctx(), initialReplica.moleculeCount, off.data(), types.data(),
                                 initialReplica.nonCrowderCount),
                "cgppd_set_system");
    if (initialReplica.has_flexible()) {
        // TorsionalLookupMatrix + the molecules' graphs (src/Molecule.cpp:30,57)
        std::vector<double> v(1600), d(1600);
        const TorsionalLookupMatrix &t = initialReplica.torsions;
        for (int a = 0; a < AA_COUNT; ++a)
            for (int b = 0; b < AA_COUNT; ++b)
                for (int n = 0; n < 4; ++n) {
                    v[(a * AA_COUNT + b) * 4 + n] = t.V[a][b][n];
                    d[(a * AA_COUNT + b) * 4 + n] = t.Sigma[a][b][n];
                }
        dev_->check(cgppd_set_torsion_table(dev_->ctx(), v.data(), d.data()), "cgppd_set_torsion_table");
        for (int m = 0; m < initialReplica.moleculeCount; ++m) {
            const Molecule &mol = initialReplica.molecules[m];
            if (!mol.is_flexible) continue;
            std::vector<int> code, chain, seg_off{0}, seg_res;
            for (const Residue &r : mol.Residues) {
                code.push_back((unsigned char)r.chainId);
                chain.push_back(r.chain_int_id);
            }
            for (const auto &sg : mol.segments) {
                seg_res.insert(seg_res.end(), sg.begin(), sg.end());
                seg_off.push_back((int)seg_res.size());
            }
            if (seg_res.empty()) seg_res.push_back(0);
            dev_->check(cgppd_set_topology(dev_->ctx(), m, code.data(), chain.data(), mol.all_flexible ? 1 : 0,
                                           (int)mol.segments.size(), seg_off.data(), seg_res.data()),
                        "cgppd_set_topology");
        }
    }
    if (parameters.repulsiveCrowding)  // crowders as u(r) = (6/r)^12 (src/Simulation.cpp:577-584)
        dev_->check(cgppd_set_repulsive_crowding(dev_->ctx(), 1), "cgppd_set_repulsive_crowding");
    const int R = (int)replica.size(), M = initialReplica.moleculeCount, N = initialReplica.residueCount;
    std::vector<float> centres, rel;
    std::vector<double> temps(R);
    std::vector<float> tsteps(R), rsteps(R);
    std::vector<int> labels(R);
    centres.reserve((size_t)R * M * 3);
    rel.reserve((size_t)R * N * 3);
    for (int r = 0; r < R; ++r) {
        const std::vector<float> c = replica[r].centres(), x = replica[r].relative_positions();
        centres.insert(centres.end(), c.begin(), c.end());
        rel.insert(rel.end(), x.begin(), x.end());
        temps[r] = replica[r].temperature;
        tsteps[r] = replica[r].translateStep;
        rsteps[r] = (float)replica[r].rotateStep;
        labels[r] = replica[r].label;
    }
    dev_->check(cgppd_set_replicas(dev_->ctx(), R, centres.data(), rel.data(), temps.data(), tsteps.data(),
                                   rsteps.data(), labels.data(), parameters.seed),
                "cgppd_set_replicas");
    for (int r = 0; r < R; ++r) replica[r].setup_device(dev_.get(), r);
}

void Simulation::initSamplingFiles()
{
    // src/Simulation.cpp:1100-1135: output/<prefix>/ with its four files
    // (appended to when present) and the sample PDBs under pdb/
    if (parameters.prefix.empty()) parameters.prefix = "cgppd_" + std::to_string(parameters.pid);
    const std::string dir = "output/" + parameters.prefix;
    ::mkdir("output", 0755);
    ::mkdir(dir.c_str(), 0755);
    ::mkdir((dir + "/pdb").c_str(), 0755);
    auto open = [&](const char *name) {
        const std::string path = dir + "/" + name;
        FILE *f = std::fopen(path.c_str(), "a+");
        if (!f) f = std::fopen(path.c_str(), "wt");
        if (!f) throw error("Could not open / create sampling file " + path);
        return f;
    };
    fractionBoundFile = open("fractionBound");
    boundConformationsFile = open("boundconformations");
    acceptanceRatioFile = open("acceptance_ratios");
    exchangeFrequencyFile = open("exchange_freq");
}

void Simulation::writeSamplingFileHeaders()
{
    // src/Simulation.cpp:1137-1162
    std::fprintf(fractionBoundFile, "# Fraction Bound\n#Iteration InstantaneousAve CumulativeAve\n");
    std::fprintf(boundConformationsFile, "# iteration; complex free energy (replica free energy); temperature;\n# "
                                         "molecule: pdb file; 1 line per other molecule in the bound state\n");
    std::fprintf(acceptanceRatioFile, "# Iteration AcceptanceRatio\n");
    std::fprintf(exchangeFrequencyFile, "# Iteration continous_ratio, inst_ratio\n");
    std::fprintf(fractionBoundFile, "Iteration:  ");
    std::fprintf(acceptanceRatioFile, "Iteration:  ");
    for (const auto &t : position_of_temperature) {
        std::fprintf(fractionBoundFile, "%0.1fKi %0.1fKc ", t.first, t.first);
        std::fprintf(acceptanceRatioFile, "%0.1fKi %0.1fKc ", t.first, t.first);
    }
    std::fprintf(fractionBoundFile, " \n");
    std::fprintf(acceptanceRatioFile, "\n");
    for (FILE *f : {fractionBoundFile, boundConformationsFile, acceptanceRatioFile, exchangeFrequencyFile})
        std::fflush(f);
}

// src/Simulation.cpp:1181-1194: the molecule files of the run, in order
void Simulation::writeFileIndex()
{
    const std::string path = "output/" + parameters.prefix + "/fileindex";
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw error("cannot create " + path);
    std::fprintf(f, "index molecule_file_path crowder(Y/N)\n");
    for (size_t i = 0; i < parameters.mdata.size(); ++i)
        std::fprintf(f, "%2d %s %s", (int)i, parameters.mdata[i].pdbfilename.c_str(),
                     parameters.mdata[i].crowder ? "Y\n" : "N\n");
    std::fclose(f);
}

void Simulation::closeSamplingFiles()
{
    for (FILE **f : {&fractionBoundFile, &acceptanceRatioFile, &exchangeFrequencyFile, &boundConformationsFile}) {
        if (*f) std::fclose(*f);
        *f = nullptr;
    }
}

void Simulation::sync_from_device()
{
    // copy energies, temperatures, labels, move statistics and coordinates
    // back into the Replica objects
    const int R = (int)replica.size(), M = initialReplica.moleculeCount, N = initialReplica.residueCount;
    std::vector<double> e(R), t(R);
    std::vector<int> lab(R);
    std::vector<int64_t> cnt((size_t)3 * R), st(R), ex(2);
    dev_->check(cgppd_get_replica_stats(dev_->ctx(), e.data(), t.data(), lab.data(), cnt.data(), st.data(), ex.data()),
                "cgppd_get_replica_stats");
    std::vector<float> pos((size_t)N * 3), rel((size_t)N * 3), cen((size_t)M * 3);
    for (int r = 0; r < R; ++r) {
        Replica &rp = replica[r];
        rp.potential = e[r];
        rp.temperature = (float)t[r];
        rp.label = (short)lab[r];
        rp.accept = (long)cnt[3 * r];
        rp.acceptA = (long)cnt[3 * r + 1];
        rp.reject = (long)cnt[3 * r + 2];
        dev_->check(cgppd_get_coordinates(dev_->ctx(), r, pos.data(), rel.data(), cen.data()), "cgppd_get_coordinates");
        int k = 0;
        for (int m = 0; m < M; ++m) {
            Molecule &mol = rp.molecules[m];
            mol.center = Vector3f{cen[3 * m], cen[3 * m + 1], cen[3 * m + 2]};
            for (Residue &res : mol.Residues) {
                res.position = Vector3f{pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
                res.relativePosition = Vector3f{rel[3 * k], rel[3 * k + 1], rel[3 * k + 2]};
                ++k;
            }
        }
    }
    totalTests = (int)ex[0];
    totalExchanges = (int)ex[1];
}

void Simulation::sync_counters()
{
    // the device counts accept / acceptA / reject per replica since the
    // start (swapped with the temperatures on exchange); the replicas hold
    // the counts since their last acceptance_ratio line, as in the reference
    const int R = (int)replica.size();
    std::vector<int64_t> cnt((size_t)3 * R);
    dev_->check(cgppd_get_replica_stats(dev_->ctx(), nullptr, nullptr, nullptr, cnt.data(), nullptr, nullptr),
                "cgppd_get_replica_stats");
    if (seen_counters_.size() != cnt.size()) seen_counters_.assign(cnt.size(), 0);
    for (int r = 0; r < R; ++r) {
        replica[r].accept += (long)(cnt[3 * r] - seen_counters_[3 * r]);
        replica[r].acceptA += (long)(cnt[3 * r + 1] - seen_counters_[3 * r + 1]);
        replica[r].reject += (long)(cnt[3 * r + 2] - seen_counters_[3 * r + 2]);
    }
    seen_counters_ = cnt;
}

void Simulation::sample_block(int step)
{
    // Replica::sample of every replica (src/Replica.cpp:1037-1093): the
    // non-crowder energy is 0 for a lone molecule, EonDeviceNC with
    // crowders or flexible molecules (unless rigid-only), else the
    // replica's running energy
    const int R = (int)replica.size(), M = initialReplica.moleculeCount, N = initialReplica.residueCount;
    std::vector<double> e(R), nc(R);
    dev_->check(cgppd_get_replica_stats(dev_->ctx(), e.data(), nullptr, nullptr, nullptr, nullptr, nullptr),
                "cgppd_get_replica_stats");
    const bool use_nc = M != initialReplica.nonCrowderCount ||
                        (initialReplica.has_flexible() && !initialReplica.calculate_rigid_potential_only);
    if (M > 1 && use_nc) dev_->check(cgppd_energy_nc_all(dev_->ctx(), nc.data()), "cgppd_energy_nc_all");
    std::vector<float> pos((size_t)N * 3);
    for (int r = 0; r < R; ++r) {
        Replica &rp = replica[r];
        rp.potential = e[r];
        rp.nonCrowderPotential = M == 1 ? 0.0f : (float)(use_nc ? nc[r] : e[r]);
        dev_->check(cgppd_get_coordinates(dev_->ctx(), r, pos.data(), nullptr, nullptr), "cgppd_get_coordinates");
        int k = 0;
        for (Molecule &mol : rp.molecules)
            for (Residue &res : mol.Residues) {
                res.position = Vector3f{pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
                ++k;
            }
        const std::string record = rp.sample(parameters.prefix.c_str(), step, BOUND_ENERGY_VALUE);
        std::fputs(record.c_str(), boundConformationsFile);
    }
    std::fflush(boundConformationsFile);
}

void Simulation::print_interval_statistics(int iteration)
{
    // src/Simulation.cpp:749-768 and 837-850: one column pair per
    // temperature, ascending
    sync_counters();
    std::fprintf(fractionBoundFile, "%9d: ", iteration);
    std::fprintf(acceptanceRatioFile, "%9d: ", iteration);
    for (const auto &t : temperature) {
        Replica &rp = replica[position_of_temperature[t.second]];
        rp.fraction_bound(fractionBoundFile);
        rp.acceptance_ratio(acceptanceRatioFile);
    }
    std::fprintf(fractionBoundFile, "\n");
    std::fflush(fractionBoundFile);
    std::fprintf(acceptanceRatioFile, "\n");
    std::fflush(acceptanceRatioFile);
}

void Simulation::exchange_on_device()
{
    // one exchange round on the device (src/Simulation.cpp:770-795); the
    // replicas whose temperatures it swapped swap their bookkeeping on the
    // host as Replica::exchangeReplicas does (src/Replica.cpp:188-210)
    const int R = (int)replica.size();
    sync_counters();
    std::vector<int64_t> before(2), after(2);
    dev_->check(cgppd_get_replica_stats(dev_->ctx(), nullptr, nullptr, nullptr, nullptr, nullptr, before.data()),
                "cgppd_get_replica_stats");
    dev_->check(cgppd_exchange(dev_->ctx()), "cgppd_exchange");
    std::vector<double> t(R);
    dev_->check(cgppd_get_replica_stats(dev_->ctx(), nullptr, t.data(), nullptr, nullptr, nullptr, after.data()),
                "cgppd_get_replica_stats");
    for (int i = 0; i < R; ++i) {
        if ((float)t[i] == replica[i].temperature) continue;
        for (int j = i + 1; j < R; ++j)
            if ((float)t[j] == replica[i].temperature && (float)t[i] == replica[j].temperature) {
                replica[i].exchangeReplicas(replica[j]);
                for (int c = 0; c < 3; ++c) std::swap(seen_counters_[3 * i + c], seen_counters_[3 * j + c]);
                position_of_temperature[replica[i].temperature] = i;
                position_of_temperature[replica[j].temperature] = j;
                break;
            }
    }
    tests = (int)(after[0] - before[0]);
    exchanges = (int)(after[1] - before[1]);
}

void Simulation::exchange_frequency()
{
    // src/Simulation.cpp:876-888
    totalExchanges += exchanges;
    totalTests += tests;
    exchangeFraction = float(exchanges) / float(tests);
    accumulativeExchangeFraction = float(totalExchanges) / float(totalTests);
    std::fprintf(exchangeFrequencyFile, "%10d %6.4f %6.4f\n", steps, accumulativeExchangeFraction, exchangeFraction);
    std::fflush(exchangeFrequencyFile);
    tests = 0;
    exchanges = 0;
}

void Simulation::run()
{
    // src/Simulation.cpp:709-873 with the MC loop of the reference's default
    // GPU build (CUDA with streams, src/Simulation.cpp:1041-1094): every
    // replica takes MCsteps / REsteps steps per exchange interval; after a
    // step t with t % samplefrequency == samplefrequency - 1 and
    // t >= sampleafter every replica is sampled, labelled
    // t - samplefrequency + 1; between intervals the fraction-bound and
    // acceptance lines are written and one exchange round is attempted
    // (REsteps - 1 of them); the last lines are written at MCsteps.  The MC
    // steps and the exchanges run on the device.
    if (!dev_) setup_device(parameters.gpuoffset);
    initSamplingFiles();
    writeFileIndex();
    writeSamplingFileHeaders();
    dev_->check(cgppd_mc_prepare(dev_->ctx()), "cgppd_mc_prepare");
    const int R = (int)replica.size();
    const int per = parameters.MCsteps / std::max(1, parameters.REsteps);
    const int sf = parameters.sampleFrequency;
    seen_counters_.assign((size_t)3 * R, 0);
    totalExchanges = totalTests = exchanges = tests = 0;
    long mcstep = 0;
    steps = 0;
    for (int interval = 0; interval < parameters.REsteps; ++interval) {
        const long end = mcstep + per;
        while (mcstep < end) {
            // the next step after which the replicas are sampled
            const long next = sf > 0 ? mcstep + ((sf - 1 - mcstep % sf) % sf) : end;
            if (next >= end) {
                dev_->check(cgppd_mc_steps(dev_->ctx(), (int)(end - mcstep)), "cgppd_mc_steps");
                mcstep = end;
                break;
            }
            dev_->check(cgppd_mc_steps(dev_->ctx(), (int)(next + 1 - mcstep)), "cgppd_mc_steps");
            mcstep = next + 1;
            if (next >= parameters.sampleStartsAfter) sample_block((int)(next - sf + 1));
        }
        if (interval + 1 < parameters.REsteps) {
            ++steps;
            if (per * (steps - 1) >= parameters.sampleStartsAfter) print_interval_statistics(per * steps);
            exchange_on_device();
            exchange_frequency();
        }
    }
    print_interval_statistics(parameters.MCsteps);
    sync_from_device();
    closeSamplingFiles();
}

}  // namespace cgppd
