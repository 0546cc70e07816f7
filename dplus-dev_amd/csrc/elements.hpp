// elements.hpp -- element tables of the amplitude builder.
//
// X-ray form factors: Cromer-Mann 4 Gaussians + c (International Tables Vol IV
// Table 2.2B) as tabulated in PDBReaderLib/PDBReaderLib.cpp:2270-2293; dummy-atom
// radii (nm) of radius type 4 (PDBReaderLib.cpp:204-216); atomic weights for
// centring on the centre of mass (PDBReaderLib.cpp:77-93).
#pragma once

#include <cstring>
#include <string>

namespace dplus {

struct ElementInfo {
    const char *symbol;
    double coef9[9];  // a1 b1 a2 b2 a3 b3 a4 b4 c
    double svg_radius;
    double weight;
};

static const ElementInfo kElements[] = {
    {"H", {0.49300, 10.51090, 0.32290, 26.1257, 0.14020, 3.14240, 0.04080, 57.79980, 0.0030}, 0.107, 1.008},
    {"C", {2.31000, 20.8439, 1.02000, 10.2075, 1.58860, 0.56870, 0.86500, 51.65120, 0.2156}, 0.1577, 12.011},
    {"N", {12.2126, 0.00570, 3.13220, 9.89330, 2.01250, 28.9975, 1.16630, 0.58260, -11.5290}, 0.08414, 14.007},
    {"O", {3.04850, 13.2771, 2.28680, 5.70110, 1.54630, 0.32390, 0.86700, 32.90890, 0.2508}, 0.130, 15.999},
    {"P", {6.43450, 1.90670, 4.17910, 27.1570, 1.78000, 0.52600, 1.49080, 68.16450, 1.1149}, 0.111, 30.97376},
    {"S", {6.90530, 1.46790, 5.20340, 22.2151, 1.43790, 0.25360, 1.58630, 56.1720, 0.86690}, 0.168, 32.06},
};
constexpr int kNumElements = sizeof(kElements) / sizeof(kElements[0]);

inline int element_index(const std::string &sym) {
    for (int i = 0; i < kNumElements; i++)
        if (sym == kElements[i].symbol) return i;
    return -1;
}

}  // namespace dplus
