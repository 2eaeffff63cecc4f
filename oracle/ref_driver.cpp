// ref_driver.cpp -- golden-vector generator built against the reference's
// own energy translation units (test infrastructure only).
//
// oracle/Makefile compiles this file together with the reference sources
// src/Potential.cpp, src/Residue.cpp, src/AminoAcids.cpp, src/AminoAcid.cpp,
// src/vector3f.cpp and src/Quaternion.cpp where they lie under the reference
// tree, into oracle/_ref/ref_driver.  Nothing from it ships.
//
// The double loop over molecule pairs below restates Replica::E
// (src/Replica.cpp:432-493) around the reference's own Residue::distance,
// calculate_LJ, calculate_DH and Potential, which are linked unchanged.
//
// Usage:
//   ref_driver consts
//       prints RT_to_kcalmol, DH_CONVERSION_FACTOR and r0_constant.
//   ref_driver energy <AminoAcids file> <pair table file> < system
//       system = "box L\nnmol M\n" then M lines "count" then N lines "aa x y z"
//       with aa the 3-letter residue name; prints LJ, DH (raw) and total,
//       then the molecule-pair matrix of totals.
//   ref_driver rotate < lines "angle ax ay az vx vy vz"
//       prints Quaternion(angle, axis).normalize().rotateVector(v) per line.
//   ref_driver normalize < lines "x y z"
//       prints Vector3f(x,y,z).normalize() per line.
#include "AminoAcids.h"
#include "Potential.h"
#include "Quaternion.h"
#include "Residue.h"
#include "vector3f.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static void print_consts()
{
    float rt = RT_to_kcalmol;
    float dhc = DH_CONVERSION_FACTOR;
    float r0 = r0_constant;
    std::printf("%.9g %.9g %.9g\n", rt, dhc, r0);
}

static int energy(const char *aa_file, const char *pair_file)
{
    AminoAcids aa;
    aa.init(aa_file, pair_file);
    float box = 0.0f;
    int nmol = 0;
    char key[64];
    if (std::scanf("%63s %f", key, &box) != 2 || std::strcmp(key, "box") != 0) return 2;
    if (std::scanf("%63s %d", key, &nmol) != 2 || std::strcmp(key, "nmol") != 0) return 2;
    std::vector<int> counts(nmol);
    for (int m = 0; m < nmol; ++m)
        if (std::scanf("%d", &counts[m]) != 1) return 2;
    std::vector<std::vector<Residue> > mols(nmol);
    for (int m = 0; m < nmol; ++m) {
        for (int i = 0; i < counts[m]; ++i) {
            char name[8];
            float x, y, z;
            if (std::scanf("%7s %f %f %f", name, &x, &y, &z) != 4) return 2;
            Residue R;
            // as Molecule::initFromPDB (src/Molecule.cpp:553-557) fills it
            R.aminoAcidIndex = aa.getAminoAcidIndex(name);
            R.electrostaticCharge = float(aa.get(R.aminoAcidIndex).electrostaticCharge);
            R.vanderWaalRadius = aa.data[R.aminoAcidIndex].vanderWaalRadius;
            R.position = Vector3f(x, y, z);
            mols[m].push_back(R);
        }
    }
    Potential total;
    std::vector<double> mat(nmol * nmol, 0.0);
    for (int I = 0; I < nmol; ++I) {
        for (int J = I + 1; J < nmol; ++J) {
            Potential p;
            for (size_t i = 0; i < mols[I].size(); ++i) {
                for (size_t j = 0; j < mols[J].size(); ++j) {
                    double r(mols[I][i].distance(mols[J][j], box) + EPS);
                    double lj = calculate_LJ(mols[I][i], mols[J][j], r, aa);
                    double dh = calculate_DH(mols[I][i], mols[J][j], r);
                    p.increment_LJ(lj);
                    p.increment_DH(dh);
                    total.increment_LJ(lj);
                    total.increment_DH(dh);
                }
            }
            mat[I * nmol + J] = mat[J * nmol + I] = p.total();
        }
    }
    std::printf("%.17g %.17g %.17g\n", total.LJ, total.DH, total.total());
    for (int I = 0; I < nmol; ++I) {
        for (int J = 0; J < nmol; ++J) std::printf("%s%.17g", J ? " " : "", mat[I * nmol + J]);
        std::printf("\n");
    }
    return 0;
}

static int rotate()
{
    double angle, ax, ay, az;
    float vx, vy, vz;
    while (std::scanf("%lf %lf %lf %lf %f %f %f", &angle, &ax, &ay, &az, &vx, &vy, &vz) == 7) {
        Vector3double axis(ax, ay, az);
        axis.normalizeInPlace();
        Quaternion q(angle, axis);
        q.normalize();
        Vector3f out = q.rotateVector(Vector3f(vx, vy, vz));
        std::printf("%.9g %.9g %.9g\n", out.x, out.y, out.z);
    }
    return 0;
}

static int normalize()
{
    float x, y, z;
    while (std::scanf("%f %f %f", &x, &y, &z) == 3) {
        Vector3f n = Vector3f(x, y, z).normalize();
        std::printf("%.9g %.9g %.9g\n", n.x, n.y, n.z);
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 1;
    std::string mode(argv[1]);
    if (mode == "consts") { print_consts(); return 0; }
    if (mode == "energy" && argc >= 4) return energy(argv[2], argv[3]);
    if (mode == "rotate") return rotate();
    if (mode == "normalize") return normalize();
    return 1;
}
