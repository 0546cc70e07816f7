// elements.hpp -- scattering species of the amplitude builder: atoms, the ions
// of those atoms that biological structures carry, and the implicit-hydrogen
// groups (CH, CH2, CH3, NH, NH2, NH3, OH, SH).  An ion has its own form factor
// and its element's radii and weight (the reference indexes those by element).
//
// X-ray form factors: Cromer-Mann 4 Gaussians + c (International Tables Vol IV
// Table 2.2B) for atoms and the fitted group coefficients, as tabulated in
// PDBReaderLib/PDBReaderLib.cpp:2270-2489; dummy-atom radii (nm) of radius types
// 4 and 5 (PDBReaderLib.cpp:204-264); van der Waals, empirical and calculated
// radii (nm) of radius types 1, 2 and 3 (PDBReaderLib.cpp:77-111; a group takes
// its heavy atom's radius, since the reference adds the radius of entry 0, which
// is zero, per hydrogen, :213-267); atomic weights for centring on the centre of
// mass (PDBReaderLib.cpp:77-264; likewise a group weighs what its heavy atom weighs).
#pragma once

#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace dplus {

struct ElementInfo {
    const char *key;    // lower-case PDB element field (columns 77-80) or group name
    const char *symbol; // display name
    double coef9[9];    // a1 b1 a2 b2 a3 b3 a4 b4 c
    double svg_radius;
    double weight;
    double vdw_radius, emp_radius, calc_radius;
};

static const ElementInfo kElements[] = {
    {" h  ", "H", {0.49300, 10.51090, 0.32290, 26.1257, 0.14020, 3.14240, 0.04080, 57.79980, 0.0030}, 0.107, 1.008, 0.12, 0.025, 0.053},
    {" c  ", "C", {2.31000, 20.8439, 1.02000, 10.2075, 1.58860, 0.56870, 0.86500, 51.65120, 0.2156}, 0.1577, 12.011, 0.17, 0.07, 0.067},
    {" n  ", "N", {12.2126, 0.00570, 3.13220, 9.89330, 2.01250, 28.9975, 1.16630, 0.58260, -11.5290}, 0.08414, 14.007, 0.155, 0.065, 0.056},
    {" o  ", "O", {3.04850, 13.2771, 2.28680, 5.70110, 1.54630, 0.32390, 0.86700, 32.90890, 0.2508}, 0.130, 15.999, 0.152, 0.06, 0.048},
    {" o1-", "O1-", {4.19160, 12.8573, 1.63969, 4.17236, 1.52673, 47.0179, -20.307, -0.01404, 21.9412}, 0.130, 15.999, 0.152, 0.06, 0.048},
    {" p  ", "P", {6.43450, 1.90670, 4.17910, 27.1570, 1.78000, 0.52600, 1.49080, 68.16450, 1.1149}, 0.111, 30.97376, 0.18, 0.1, 0.098},
    {" s  ", "S", {6.90530, 1.46790, 5.20340, 22.2151, 1.43790, 0.25360, 1.58630, 56.1720, 0.86690}, 0.168, 32.06, 0.18, 0.1, 0.088},
    {"na  ", "Na", {4.76260, 3.28500, 3.17360, 8.84220, 1.26740, 0.31360, 1.11280, 129.42400, 0.676}, 0.180, 22.98976, 0.227, 0.18, 0.19},
    {"na1+", "Na1+", {3.25650, 2.66710, 3.93620, 6.11530, 1.39980, 0.20010, 1.00320, 14.03900, 0.404}, 0.180, 22.98976, 0.227, 0.18, 0.19},
    {"mg  ", "Mg", {5.42040, 2.82750, 2.17350, 79.2611, 1.22690, 0.38080, 2.30730, 7.19370, 0.8584}, 0.160, 24.3050, 0.173, 0.15, 0.145},
    {"mg2+", "Mg2+", {3.49880, 2.16760, 3.83780, 4.75420, 1.32840, 0.18500, 0.84970, 10.14110, 0.4853}, 0.160, 24.3050, 0.173, 0.15, 0.145},
    {"cl  ", "Cl", {11.46040, 0.01040, 7.19640, 1.16620, 6.25560, 18.5194, 1.64550, 47.77840, -9.5574}, 0.100, 35.45, 0.175, 0.1, 0.079},
    {"cl1-", "Cl1-", {18.29150, 0.00660, 7.40840, 1.17170, 6.53370, 19.5424, 2.33860, 60.44860, -16.378}, 0.100, 35.45, 0.175, 0.1, 0.079},
    {" k  ", "K", {8.21860, 12.79490, 7.43980, 0.77480, 1.05190, 213.187, 0.86590, 41.68410, 1.4228}, 0.220, 39.0983, 0.275, 0.22, 0.243},
    {" k1+", "K1+", {7.95780, 12.63310, 7.49170, 0.76740, 6.35900, -0.0020, 1.19150, 31.91280, -4.9978}, 0.220, 39.0983, 0.275, 0.22, 0.243},
    {"ca  ", "Ca", {8.62660, 10.44210, 7.38730, 0.65990, 1.58990, 85.7484, 1.02110, 178.43700, 1.3751}, 0.197, 40.078, 0.231, 0.18, 0.194},
    {"ca2+", "Ca2+", {15.63480, -0.00740, 7.95180, 0.60890, 8.43720, 10.3116, 0.85370, 25.99050, -14.875}, 0.197, 40.078, 0.231, 0.18, 0.194},
    {"mn  ", "Mn", {11.28190, 5.34090, 7.35730, 0.34320, 3.01930, 17.86740, 2.24410, 83.75430, 1.0896}, 0.130, 54.93804, 0.186, 0.14, 0.161},
    {"mn2+", "Mn2+", {10.80610, 5.27960, 7.36200, 0.34350, 3.52680, 14.34300, 0.21840, 41.32350, 1.0874}, 0.130, 54.93804, 0.186, 0.14, 0.161},
    {"fe  ", "Fe", {11.76950, 4.76110, 7.35730, 0.30720, 3.52220, 15.35350, 2.30450, 76.88050, 1.0369}, 0.124, 55.845, 0.181, 0.14, 0.156},
    {"fe2+", "Fe2+", {11.04240, 4.65380, 7.37400, 0.30530, 4.13460, 12.05460, 0.43990, 31.28090, 1.0097}, 0.124, 55.845, 0.181, 0.14, 0.156},
    {"fe3+", "Fe3+", {11.17640, 4.61470, 7.38630, 0.30050, 3.39480, 11.67290, 0.07240, 38.55660, 0.9707}, 0.124, 55.845, 0.181, 0.14, 0.156},
    {"cu  ", "Cu", {13.33800, 3.58280, 7.16760, 0.24700, 5.61580, 11.39660, 1.67350, 64.81260, 1.191}, 0.128, 63.546, 0.14, 0.135, 0.145},
    {"cu1+", "Cu1+", {11.94750, 3.36690, 7.35730, 0.22740, 6.24550, 8.66250, 1.55780, 25.84870, 0.89}, 0.128, 63.546, 0.14, 0.135, 0.145},
    // b2 as the reference tabulates it (International Tables: 0.244078)
    {"cu2+", "Cu2+", {11.81680, 3.37484, 7.11181, 244078, 5.78135, 7.98760, 1.14523, 19.89700, 1.14431}, 0.128, 63.546, 0.14, 0.135, 0.145},
    {"zn  ", "Zn", {14.07430, 3.26550, 7.03180, 0.23330, 5.16520, 10.31630, 2.41000, 58.70970, 1.3041}, 0.133, 65.38, 0.139, 0.135, 0.142},
    {"zn2+", "Zn2+", {11.97190, 2.99460, 7.38620, 0.20310, 6.46680, 7.08260, 1.39400, 18.09950, 0.7807}, 0.133, 65.38, 0.139, 0.135, 0.142},
    // implicit-hydrogen groups
    {"ch", "CH", {0.894937, 55.7145, 0.894429, 4.03158, 3.78824, 24.8323, 3.14683e-6, 956.628, 1.42149}, 0.173, 12.011, 0.17, 0.07, 0.067},
    {"ch2", "CH2", {1.61908, 52.1451, 2.27205, 24.6589, 2.1815, 24.6587, 0.0019254, 152.165, 1.92445}, 0.185, 12.011, 0.17, 0.07, 0.067},
    {"ch3", "CH3", {12.5735, 38.7341, -0.456658, -6.28167, 5.71547, 54.955, -11.711, 47.898, 2.87762}, 0.197, 12.011, 0.17, 0.07, 0.067},
    {"nh", "NH", {0.00506991, 108.256, 2.03147, 14.6199, 1.82122, 14.628, 2.06506, 35.4102, 2.07168}, 0.122, 14.007, 0.155, 0.065, 0.056},
    {"nh2", "NH2", {3.00872, 28.3717, 0.288137, 63.9637, 3.39248, 3.51866, 2.03511, 28.3675, 0.269952}, 0.145, 14.007, 0.155, 0.065, 0.056},
    {"nh3", "NH3", {0.294613, 67.4408, 6.48379, 29.1576, 5.67182, 0.54735, 6.57164, 0.547493, -9.02757}, 0.162, 14.007, 0.155, 0.065, 0.056},
    {"oh", "OH", {-2.73406, 22.1288, 0.00966263, 94.3428, 6.64439, 13.9044, 2.67949, 32.7607, 2.39981}, 0.150, 15.999, 0.152, 0.06, 0.048},
    {"sh", "SH", {-127.811, 7.19935, 62.5514, 12.1591, 160.747, 1.88979, 2.34822, 55.952, -80.836}, 0.181, 32.06, 0.18, 0.1, 0.088},
};
constexpr int kNumElements = sizeof(kElements) / sizeof(kElements[0]);

// species index of a lower-case key (" c  ", "mg  ", "ch2", ...), or -1.  A
// charge may follow the element in either order ("o1-", "o-1"); charged
// nitrogen takes neutral nitrogen's form factor, as in the reference's ion map
// (getAtomIonIndices, PDBReaderLib.cpp:1971-2205)
inline int species_index(std::string key) {
    if (key.size() == 4 && (key[2] == '+' || key[2] == '-') && key[3] >= '0' && key[3] <= '9') std::swap(key[2], key[3]);
    if (key == " n1+") key = " n  ";
    for (int i = 0; i < kNumElements; i++)
        if (key == kElements[i].key) return i;
    return -1;
}

// species index of an element symbol ("C", "Mg", "CH2" ...), or -1
inline int element_index(const std::string &sym) {
    for (int i = 0; i < kNumElements; i++)
        if (sym == kElements[i].symbol) return i;
    return -1;
}

// atoms of a PDB file: coordinates (nm) and species (index into kElements)
struct PdbAtoms {
    std::vector<double> xyz;
    std::vector<int> species;
    int implicit_atoms = 0, implicit_residues = 0;
    // the record fields the reference keeps for writing the model back out
    // (PDBReaderLib.cpp:862-931): serial (cols 7-11), name (13-16), residue
    // name (18-20), chain (22), residue number (23-26), occupancy (55-60),
    // temperature factor (61-66), segment (73-76), element field (77-80)
    std::vector<std::string> serial, name, res_name, res_no, seg, elem4;
    std::vector<char> chain;
    std::vector<double> occ, bfac;
};
PdbAtoms read_pdb(const std::string &path);

}  // namespace dplus
