// cgppd_host.h -- C++ host side of the MI355X REMC path, keeping the names
// and semantics of the reference's Replica / Simulation / AminoAcids /
// Potential / Molecule API (src/Replica.h:18-213, src/Simulation.h:31-109,
// src/AminoAcids.h:8-27, src/Potential.h:19-92, src/Molecule.h,
// src/structures.h:49-127) over the C ABI of libcgppd_hip.so.
//
// What is kept: class and method names, the argfile / CLI grammar, the data
// formats, Replica::E() as the CPU energy, EonDevice / EonDeviceNC /
// EonDeviceAsync + SumGridResults / ReplicaDataToDevice /
// MoleculeDataToDevice as device entry points, the temperature and step
// ladders, and the exchange bookkeeping.  What changes: one DeviceContext
// (one GPU, one stream) holds all replicas; Simulation::run drives every
// replica's MC steps and the exchanges on the device (cgppd_run) instead of
// host threads and per-step host round trips.
#ifndef CGPPD_HOST_H
#define CGPPD_HOST_H

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cgppd_hip.h"

namespace cgppd {

constexpr int AA_COUNT = 20;

// ---------------------------------------------------------------------------
// Constants (src/constants.h:8-46, src/definitions.h:10-44), float as there.
// ---------------------------------------------------------------------------
float rt_to_kcalmol();
float dh_conversion_factor();
float r0_constant();
constexpr float EPS = 1e-38f;
constexpr float BOUND_ENERGY_VALUE = -1.1844f;
constexpr float MIN_TRANSLATION = 0.5f, MAX_TRANSLATION = 2.0f;
constexpr float MIN_ROTATION = 0.2f, MAX_ROTATION = 0.4f;

struct Vector3f {
    float x = 0.f, y = 0.f, z = 0.f;
};

// ---------------------------------------------------------------------------
// Amino-acid data and the contact-energy table (src/AminoAcids.cpp:19-194).
// ---------------------------------------------------------------------------
struct AminoAcid {
    std::string name, sname;
    char shortSymbol = ' ';
    float vanderWaalRadius = 0.f;
    float electrostaticCharge = 0.f;
    int index = 0;
    // src/AminoAcid.cpp:19-25: the three-letter name in upper case
    std::string getSNAME() const
    {
        std::string u = sname;
        for (char &c : u) c = (char)std::toupper((unsigned char)c);
        return u;
    }
};

class AminoAcids {
public:
    AminoAcid data[AA_COUNT];
    float LJpotentials[AA_COUNT][AA_COUNT];
    bool clean = false;

    // throws std::runtime_error when a file is missing or malformed
    void init(const char *aminoAcidFile, const char *pairPotentialFile);
    bool loadAminoAcidData(const char *filename);
    bool loadLJPotentialData(const char *filename);
    int getAminoAcidIndex(const char *name) const;
    const AminoAcid &get(int i) const { return data[i]; }
};

// ---------------------------------------------------------------------------
// Potential accumulator (src/Potential.h:19-92) of the flexible-linker build:
// raw LJ [RT], raw DH, raw bond sum sum (r - R0)^2, the product of the angle
// terms and the torsion sum; rigid systems leave the last three at 0, 1, 0.
// ---------------------------------------------------------------------------
class Potential {
public:
    double LJ = 0.0, DH = 0.0, bond = 0.0, angle = 1.0, torsion = 0.0;
    Potential() = default;
    Potential(double LJ_kcal, double DH_kcal);  // stores raw, as src/Potential.cpp:116-126
    Potential(double LJ_kcal, double DH_kcal, double bond_kcal, double angle_kcal, double torsion_kcal);
    void increment_LJ(double v) { LJ += v; }
    void increment_DH(double v) { DH += v; }
    void increment_bond(double v) { bond += v; }
    void increment_angle(double v) { angle *= v; }
    void increment_torsion(double v) { torsion += v; }
    Potential &operator+=(const Potential &o)
    {
        LJ += o.LJ;
        DH += o.DH;
        bond += o.bond;
        angle *= o.angle;
        torsion += o.torsion;
        return *this;
    }
    double total_LJ() const;
    double total_DH() const;
    double total_bond() const;
    double total_angle() const;
    double total_torsion() const;
    double total() const;
    bool almost_equal(const Potential &p, float eps = 0.00001f) const;
};

// Torsion table (src/TorsionalLookupMatrix.cpp:21-52): V [kcal/mol] and the
// phase [rad] for the middle residue types of a torsion, n = 1..4.
class TorsionalLookupMatrix {
public:
    double V[AA_COUNT][AA_COUNT][4] = {};
    double Sigma[AA_COUNT][AA_COUNT][4] = {};
    bool loadData(const char *filename, const AminoAcids &aminoAcids);  // throws on a missing file
    double getV(int ai, int aj, int n) const { return V[ai][aj][n - 1]; }
    double getSigma(int ai, int aj, int n) const { return Sigma[ai][aj][n - 1]; }
};

// CPU pair terms with the reference's rounding (src/Potential.cpp:5-42).
double calculate_LJ(int ti, int tj, double r, const AminoAcids &aa);
double calculate_DH(float qi, float qj, double r);

// ---------------------------------------------------------------------------
// Residues and rigid molecules (src/Residue.h, src/Molecule.cpp:25-266,
// 506-612).
// ---------------------------------------------------------------------------
struct Residue {
    int aminoAcidIndex = -2;
    float electrostaticCharge = 0.f;
    float vanderWaalRadius = 0.f;
    Vector3f position, relativePosition;
    char chainId = ' ';
    int chain_int_id = 0;
    int resSeq = -1;
    int moleculeId = -1;
    // replica-wide UIDs (src/Geometry.cpp:256-302), 0 = none
    int chain_UID = 0, rigid_domain_UID = 0, segment_bond_UID = 0;
    // src/Residue.cpp:62-82: FP32 minimum image, sqrtf
    double distance(const Residue &r, float bounding_value = 0.0f) const;
    // src/Residue.cpp:116-129
    bool chain_neighbour_of(const Residue &o) const { return chain_UID == o.chain_UID && std::abs(resSeq - o.resSeq) < 4; }
    bool same_rigid_domain_as(const Residue &o) const { return rigid_domain_UID && rigid_domain_UID == o.rigid_domain_UID; }
    bool bonded_to(const Residue &o) const { return segment_bond_UID && segment_bond_UID == o.segment_bond_UID; }
};

struct moldata {
    std::string pdbfilename, name;
    float px = 0, py = 0, pz = 0;
    bool translate = false;
    float rx = 0, ry = 0, rz = 0, ra = 0;
    bool crowder = false;
    bool all_flexible = false;
    std::vector<std::vector<int>> segments;
};

class Molecule {
public:
    std::vector<Residue> Residues;
    int residueCount = 0;
    int chainCount = 0;
    int index = -2;
    Vector3f center;
    float bounding_value = 0.f;
    bool crowder = false;
    std::string filename, name;
    // flexible linkers: the bonded graph from the library's topology builder
    // (cgppd_topology_list, src/Geometry.cpp:34-192)
    bool is_flexible = false, all_flexible = false;
    std::vector<std::vector<int>> segments;
    std::vector<int> bonds, angles, torsions, indirect;  // 2, 3, 4, 2 residue indices each
    std::vector<int> domain_of, segbond_of;              // per residue, -1 = none
    int vertex_count = 0;
    const TorsionalLookupMatrix *torsion_data = nullptr;

    void init(const moldata &mol, const AminoAcids &aa, int index, float bounding_value,
              const TorsionalLookupMatrix *torsions = nullptr);
    void init_graph();
    // src/Molecule.cpp:636-699: intra-molecular pairs (include_LJ_and_DH) or
    // minus the indirect neighbours, plus the bonded terms
    // (tor: the torsion table; the molecule's own torsion_data when null)
    Potential E(bool include_LJ_and_DH, const AminoAcids &aa, const TorsionalLookupMatrix *tor = nullptr) const;
    void initFromPDB(const char *pdbfilename, const AminoAcids &aa);
    void setPosition(Vector3f c);                     // src/Molecule.cpp:183-191
    void translate(Vector3f v);                       // src/Molecule.cpp:146-159
    void rotate(const double axis[3], double angle);  // src/Molecule.cpp:193-208
};

// ---------------------------------------------------------------------------
// Arguments (src/structures.h:78-127) and the argfile / CLI parsers
// (src/Simulation.cpp:78-453).
// ---------------------------------------------------------------------------
struct argdata {
    bool viewConditions = false;
    bool skipsimulation = false;
    int cuda_blockSize = 64;
    bool auto_blockdim = true;
    int threads = 1;
    int streams = 1;
    int max_replicas_per_thread = 0;
    int gpus = 1;
    int gpuoffset = 0;
    int MCsteps = 5000;
    int REsteps = 10;
    int replicas = 20;
    float bound = 150.0f;
    int sampleFrequency = 5000;
    int sampleStartsAfter = 5000;
    int checkpointFrequency = 10000;
    std::string prefix, title, file;
    bool inputFile = false;
    int nonCrowders = 0;
    float temperatureMin = 250.0f;
    float temperatureMax = 500.0f;
    int pid = 0;
    bool resume = false;
    uint64_t seed = 0;         // new: keys the device Philox streams
    bool use_graph = true;     // new: replay exchange intervals from a hipGraph
    bool repulsiveCrowding = false;  // new: run-time REPULSIVE_CROWDING (src/definitions.h:18)
    std::string aminoAcidFile = "data/AminoAcids";
    std::string pairPotentialFile = "data/pair_potentials_MJ1996";
    std::string torsionFile = "data/torsional_pair_potentials";  // src/definitions.h:7
    std::vector<moldata> mdata;
    std::map<std::string, int> mdata_map;
};

// ---------------------------------------------------------------------------
// Device context: replaces setup_CUDA / teardown_CUDA
// (src/Simulation.h:97-105, src/Simulation.cpp:892-948).
// ---------------------------------------------------------------------------
class DeviceContext {
public:
    DeviceContext(int device, float box_dimension, const AminoAcids &aa);
    ~DeviceContext();
    DeviceContext(const DeviceContext &) = delete;
    DeviceContext &operator=(const DeviceContext &) = delete;
    cgppd_ctx *ctx() const { return ctx_; }
    void check(int rc, const char *what) const;  // throws std::runtime_error
private:
    cgppd_ctx *ctx_ = nullptr;
};

class Replica {
public:
    AminoAcids aminoAcids;
    std::vector<Molecule> molecules;
    int moleculeCount = 0;
    int residueCount = 0;
    int nonCrowderCount = 0;
    int nonCrowderResidues = 0;
    int max_residue_count = 0;
    float boundingValue = 0.f;

    float temperature = 300.0f;
    short label = -1;
    double potential = 0.0;
    double newPotential = 0.0;
    float translateStep = 0.5f;
    double rotateStep = 0.2;
    long accept = 0, acceptA = 0, reject = 0;
    long totalAccept = 0, totalAcceptReject = 0;
    float acceptanceRatio = 0.f, accumulativeAcceptanceRatio = 0.f;
    int boundSamples = 0, samplesSinceLastExchange = 0, totalBoundSamples = 0, totalSamples = 0;
    float fractionBound = 0.f, accumulativeFractionBound = 0.f;
    float nonCrowderPotential = 0.0f;  // src/Replica.h:106
    // the reference's compile-time REPULSIVE_CROWDING as a run-time flag:
    // crowder pairs contribute (6/r)^12 below 15 A and no Debye-Hueckel term
    bool repulsiveCrowding = false;
    // src/Replica.h:55: tests skip the internal potential; every residue of
    // a molecule counts as one rigid domain on the device (src/Replica.cpp:753-757)
    bool calculate_rigid_potential_only = false;

    // src/Replica.cpp:19-98
    void init_first_replica(const argdata &parameters, const AminoAcids &amino_acid_data, int initial_molecule_array_size);
    void init_child_replica(const Replica &ir, int index, double geometricTemperature, double geometricRotation,
                            double geometricTranslate, const argdata &parameters);
    void loadMolecule(const moldata &mol);
    void copy(const Replica &r);
    void exchangeReplicas(Replica &r);  // src/Replica.cpp:188-210

    // CPU energies (src/Replica.cpp:421-560)
    Potential E() const;
    Potential E(const Molecule *a, const Molecule *b) const;
    Potential internal_molecule_E(bool include_LJ_and_DH) const;
    bool has_flexible() const;
    TorsionalLookupMatrix torsions;

    // device entry points (src/Replica.cpp:705-935); no CPU fallback
    void setup_device(DeviceContext *dev, int slot);
    void ReplicaDataToDevice();
    void MoleculeDataToDevice(int moleculeId);
    double EonDevice();
    double EonDeviceNC();
    void EonDeviceAsync();
    double SumGridResults();

    // flat views used by the batched device set-up
    std::vector<int> molecule_offsets() const;
    std::vector<int> residue_types() const;
    std::vector<float> positions() const;
    std::vector<float> relative_positions() const;
    std::vector<float> centres() const;

    // src/Replica.cpp:1037-1093: one sample of this replica at
    // current_step, its non-crowder energy already in nonCrowderPotential
    // (the Simulation evaluates every replica's at once, cgppd_energy_nc_all);
    // counts it (bound below boundEnergyThreshHold) and, as the reference
    // records every sample, writes its PDB under output/<prefix>/pdb and
    // returns the boundconformations record.
    std::string sample(const char *prefix, int current_step, float boundEnergyThreshHold);
    // src/Replica.cpp:1095-1122
    void fraction_bound(FILE *f);
    void acceptance_ratio(FILE *f);
    // src/Replica.cpp:1124-1168
    void saveAsSinglePDB(const char *filename, const char *prefix, int sample, bool skip_crowders = true) const;
    bool flexible_molecule_exists() const { return has_flexible(); }

private:
    DeviceContext *dev_ = nullptr;
    int slot_ = -1;
};

class Simulation {
public:
    argdata parameters;
    AminoAcids aminoAcidData;
    Replica initialReplica;
    std::vector<Replica> replica;
    std::map<float, int> position_of_temperature;
    std::map<int, float> temperature;
    double geometricTemperature = 1.0, geometricTranslate = 1.0, geometricRotation = 1.0;
    int exchanges = 0, tests = 0, totalExchanges = 0, totalTests = 0, offset = 0, steps = 0;
    float exchangeFraction = 0.f, accumulativeExchangeFraction = 0.f;
    float near_300k = 0.f;

    // src/Simulation.cpp:516-701 (argument passes, data, replicas, ladders)
    void init(int argc, char **argv, int pid);
    void getArgs(int argc, char **argv, bool first_pass);
    void loadArgsFromFile();
    void check_and_modify_parameters();
    void printHelp() const;

    // src/Simulation.cpp:709-873: all MC steps and exchanges on the device
    void run();

    // device set-up shared by run() and the tests
    void setup_device(int device);
    DeviceContext *device() const { return dev_.get(); }

    ~Simulation();

private:
    std::unique_ptr<DeviceContext> dev_;
    FILE *fractionBoundFile = nullptr, *acceptanceRatioFile = nullptr, *exchangeFrequencyFile = nullptr,
         *boundConformationsFile = nullptr;
    void initSamplingFiles();
    void writeSamplingFileHeaders();
    void closeSamplingFiles();
    void writeFileIndex();
    void sync_from_device();
    // the device's move counters at the last sync, per replica (they travel
    // with the temperatures on exchange, as the replicas' counters do)
    std::vector<int64_t> seen_counters_;
    void sync_counters();
    void sample_block(int step);
    void print_interval_statistics(int iteration);
    void exchange_on_device();
    void exchange_frequency();
};

}  // namespace cgppd

#endif
