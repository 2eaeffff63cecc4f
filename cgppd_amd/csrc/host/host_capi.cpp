// host_capi.cpp -- C entry points over the C++ host API, used by the Python
// test harness (the C++ classes are the drop-in; this only exposes them).
#include <cstring>
#include <string>
#include <vector>

#include "cgppd_host.h"

namespace {
thread_local std::string g_err;

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown C++ exception";
    }
    return -1;
}
}  // namespace

extern "C" {

const char *cgppd_host_last_error(void) { return g_err.c_str(); }

// AminoAcids::init: eps[400], sigma[20], charge[20]
int cgppd_host_load_force_field(const char *aa_file, const char *pair_file, float *eps, float *sigma, float *charge)
{
    return guarded([&] {
        cgppd::AminoAcids aa;
        aa.init(aa_file, pair_file);
        for (int i = 0; i < cgppd::AA_COUNT; ++i) {
            sigma[i] = aa.data[i].vanderWaalRadius;
            charge[i] = aa.data[i].electrostaticCharge;
            for (int j = 0; j < cgppd::AA_COUNT; ++j) eps[i * cgppd::AA_COUNT + j] = aa.LJpotentials[i][j];
        }
    });
}

// Simulation::init from a CLI argument vector; returns an opaque handle
int cgppd_host_simulation_create(int argc, char **argv, void **out)
{
    return guarded([&] {
        auto *sim = new cgppd::Simulation();
        try {
            sim->init(argc, argv, 0);
        } catch (...) {
            delete sim;
            throw;
        }
        *out = sim;
    });
}

int cgppd_host_simulation_destroy(void *h)
{
    return guarded([&] { delete static_cast<cgppd::Simulation *>(h); });
}

// sizes: replicas, molecules, residues
int cgppd_host_simulation_sizes(void *h, int *n_rep, int *n_mol, int *n_res)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        *n_rep = (int)s->replica.size();
        *n_mol = s->initialReplica.moleculeCount;
        *n_res = s->initialReplica.residueCount;
    });
}

// topology and the coordinates / ladders of replica r
int cgppd_host_replica_state(void *h, int r, int *mol_off, int *types, float *pos, double *temperature,
                             float *translate_step, double *rotate_step)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        const cgppd::Replica &rp = s->replica.at(r);
        const auto off = rp.molecule_offsets();
        const auto ty = rp.residue_types();
        const auto x = rp.positions();
        if (mol_off) std::memcpy(mol_off, off.data(), sizeof(int) * off.size());
        if (types) std::memcpy(types, ty.data(), sizeof(int) * ty.size());
        if (pos) std::memcpy(pos, x.data(), sizeof(float) * x.size());
        if (temperature) *temperature = rp.temperature;
        if (translate_step) *translate_step = rp.translateStep;
        if (rotate_step) *rotate_step = rp.rotateStep;
    });
}

// Replica::E() (CPU): raw LJ, DH and total
int cgppd_host_replica_E(void *h, int r, double *lj, double *dh, double *total)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        const cgppd::Potential p = s->replica.at(r).E();
        *lj = p.LJ;
        *dh = p.DH;
        *total = p.total();
    });
}

// Replica::internal_molecule_E(include) (CPU): raw LJ, DH, bond, angle
// product, torsion and the total
int cgppd_host_replica_internal(void *h, int r, int include, double *out)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        const cgppd::Potential p = s->replica.at(r).internal_molecule_E(include != 0);
        out[0] = p.LJ;
        out[1] = p.DH;
        out[2] = p.bond;
        out[3] = p.angle;
        out[4] = p.torsion;
        out[5] = p.total();
    });
}

// device set-up, then Replica::EonDevice / EonDeviceNC / async
int cgppd_host_setup_device(void *h, int device)
{
    return guarded([&] { static_cast<cgppd::Simulation *>(h)->setup_device(device); });
}

int cgppd_host_replica_EonDevice(void *h, int r, double *total, double *total_nc, double *total_async)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        cgppd::Replica &rp = s->replica.at(r);
        rp.ReplicaDataToDevice();
        *total = rp.EonDevice();
        *total_nc = rp.EonDeviceNC();
        rp.EonDeviceAsync();
        *total_async = rp.SumGridResults();
    });
}

// Replica::calculate_rigid_potential_only of replica r (takes effect at the
// next ReplicaDataToDevice, as in the reference)
int cgppd_host_replica_set_rigid_only(void *h, int r, int flag)
{
    return guarded([&] { static_cast<cgppd::Simulation *>(h)->replica.at(r).calculate_rigid_potential_only = flag != 0; });
}

// Simulation::run, then the per-replica running energies and temperatures
int cgppd_host_simulation_run(void *h, double *energy, double *temperature)
{
    return guarded([&] {
        auto *s = static_cast<cgppd::Simulation *>(h);
        s->run();
        for (size_t r = 0; r < s->replica.size(); ++r) {
            if (energy) energy[r] = s->replica[r].potential;
            if (temperature) temperature[r] = s->replica[r].temperature;
        }
    });
}

}  // extern "C"
