// pdb.cpp -- PDB text to scattering species and coordinates for the amplitude
// builder, following PDBReaderOb::readPDBstream (PDBReaderLib/PDBReaderLib.cpp:816-1075):
//  * ATOM/HETATM records of every model, alternate locations other than ' '
//    and 'A' skipped (:833-834);
//  * the element field is columns 77-80 (padded to 4 characters, or taken from
//    atom-name columns 13-14 when blank, :836-859), looked up with its first
//    two letters in lower case (getAtomIonIndices, :1971-2205);
//  * coordinates columns 31-54 in Angstrom -> nm (:872-874);
//  * a residue with no hydrogen atom whose name and atom labels are all known
//    has its atoms relabelled as implicit-hydrogen groups CH, CH2, CH3, NH,
//    NH2, NH3, OH, SH (ChangeResidueGroupsToImplicit, :1533-1758; pH 7.4).

#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "device.hpp"
#include "elements.hpp"

namespace dplus {

namespace {

// residue -> "label:group ..." (an empty group keeps the atom's element);
// amino acids, RNA and DNA nucleotides as in PDBReaderLib.cpp:1536-1720
const char *const kResidueGroups[][2] = {
    {"arg", "n:nh ca:ch c: o: cb:ch2 cg:ch2 cd:ch2 ne:nh cz: nh:nh2"},
    {"his", "n:nh ca:ch c: o: cb:ch2 cg: nd:nh cd:ch ce:ch ne:nh"},
    {"lys", "n:nh ca:ch c: o: cb:ch2 cg:ch2 cd:ch2 ce:ch2 nz:nh3"},
    {"asp", "n:nh ca:ch c: o: cb:ch2 cg: od:"},
    {"glu", "n:nh ca:ch c: o: cb:ch2 cg:ch2 cd: oe:"},
    {"ser", "n:nh ca:ch c: o: cb:ch2 og:oh"},
    {"thr", "n:nh ca:ch c: o: cb:ch cg:ch3 og:oh"},
    {"asn", "n:nh ca:ch c: o: cb:ch2 cg: od: nd:nh2"},
    {"gln", "n:nh ca:ch c: o: cb:ch2 cg:ch2 cd: oe: ne:nh2"},
    {"cys", "n:nh ca:ch c: o: cb:ch2 sg:sh"},
    {"sec", "n:nh ca:ch c: o: cb:ch2 seg:"},
    {"gly", "n:nh ca:ch2 c: o:"},
    {"pro", "n: ca:ch c: o: cb:ch2 cg:ch2 cd:ch2"},
    {"ala", "n:nh ca:ch c: o: cb:ch3"},
    {"val", "n:nh ca:ch c: o: cb:ch cg:ch3"},
    {"ile", "n:nh ca:ch c: o: cb:ch cg1:ch2 cg2:ch3 cd1:ch3 cd:ch3"},
    {"leu", "n:nh ca:ch c: o: cb:ch2 cg:ch cd:ch3"},
    {"met", "n:nh ca:ch c: o: cb:ch2 cg:ch2 sd: ce:ch3"},
    {"phe", "n:nh ca:ch c: o: cb:ch2 cg: cd:ch ce:ch cz:ch"},
    {"tyr", "n:nh ca:ch c: o: cb:ch2 cg: cd:ch ce:ch cz: oh:oh"},
    {"trp", "n:nh ca:ch c: o: cb:ch2 cg: cd1:ch cd2: ne1:nh ce2: ce3:ch cz2:ch cz3:ch ch2:ch"},
    {"a", "p: op1: op2: op: op3: o5': c5':ch2 c4':ch o4': c3':ch o3': c2':ch o2':oh c1':ch "
          "n9: c8:ch n7: c6: c5: n6:nh2 n1: c2:ch n3: c4:"},
    {"u", "p: op1: op2: op: op3: o5': c5':ch2 c4':ch o4': c3':ch o3': c2':ch o2':oh c1':ch "
          "n1: c2: o2: n3:nh c4: o4: c5:ch c6:ch"},
    {"c", "p: op1: op2: op: op3: o5': c5':ch2 c4':ch o4': c3':ch o3': c2':ch o2':oh c1':ch "
          "n1: c2: o2: n3: c4: n4:nh2 c5:ch c6:ch"},
    {"g", "p: op1: op2: op: op3: o5': c5':ch2 c4':ch o4': c3':ch o3': c2':ch o2':oh c1':ch "
          "n9: c8:ch n7: c5: c6: o6: n1:nh c2: n2:nh2 n3: c4:"},
    {"t", "p: op1: op2: op: op3: o5': c5':ch2 c4':ch o4': c3':ch o3': c2':ch o2':oh c1':ch "
          "n1: c2: o2: n3:nh c4: o4: c5: c7:ch3 c6:ch"},
    {"da", "p: op1: op2: op: op3: o5': c5':ch2 o4': c4':ch o3': c3':ch c2':ch2 c1':ch "
           "n9: c8:ch n7: c6: c5: n6:nh2 n1: c2:ch n3: c4:"},
    {"du", "p: op1: op2: op: op3: o5': c5':ch2 o4': c4':ch o3': c3':ch c2':ch2 c1':ch "
           "n1: c2: o2: n3:nh c4: o4: c5:ch c6:ch"},
    {"dc", "p: op1: op2: op: op3: o5': c5':ch2 o4': c4':ch o3': c3':ch c2':ch2 c1':ch "
           "n1: c2: o2: n3: c4: n4:nh2 c5:ch c6:ch"},
    {"dg", "p: op1: op2: op: op3: o5': c5':ch2 o4': c4':ch o3': c3':ch c2':ch2 c1':ch "
           "n9: c8:ch n7: c5: c6: o6: n1:nh c2: n2:nh2 n3: c4:"},
    {"dt", "p: op1: op2: op: op3: o5': c5':ch2 o4': c4':ch o3': c3':ch c2':ch2 c1':ch "
           "n1: c2: o2: n3:nh c4: o4: c5: c7:ch3 c6:ch"},
};

const std::map<std::string, std::map<std::string, std::string>> &residue_groups() {
    static const std::map<std::string, std::map<std::string, std::string>> table = [] {
        std::map<std::string, std::map<std::string, std::string>> t;
        for (auto &r : kResidueGroups) {
            std::istringstream in(r[1]);
            std::string item;
            auto &m = t[r[0]];
            while (in >> item) {
                const size_t c = item.find(':');
                m.emplace(item.substr(0, c), item.substr(c + 1));
            }
        }
        return t;
    }();
    return table;
}

std::string strip(const std::string &s) {
    std::string o;
    for (char c : s)
        if (!isspace((unsigned char)c)) o += c;
    return o;
}

std::string lower(std::string s) {
    for (auto &c : s) c = (char)tolower((unsigned char)c);
    return s;
}

struct Record {
    double x, y, z, occ, bfac;
    std::string elem4, name, res_name, res_no, serial, seg;
    char chain;
};

// labels of one residue as implicit groups, or nothing when any label is unknown
bool implicit_labels(const std::vector<Record> &R, size_t i0, size_t n, std::vector<std::string> &out) {
    const std::string res = lower(strip(R[i0].res_name));
    auto aa = residue_groups().find(res);
    if (aa == residue_groups().end()) return false;
    const bool keep_digits = res == "trp" || res == "ile" || res.size() < 3;
    std::vector<std::string> labels(n);
    for (size_t k = 0; k < n; k++) {
        std::string a = strip(R[i0 + k].name);
        if (!keep_digits) {
            std::string d;
            for (char c : a)
                if (!isdigit((unsigned char)c)) d += c;
            a = d;
        }
        a = lower(a);
        auto it = aa->second.find(a);
        if (a == "oxt") continue;
        if (it == aa->second.end()) return false;
        labels[k] = it->second;
    }
    out = labels;
    return true;
}

}  // namespace

PdbAtoms read_pdb(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw Error(kErrFileNotFound, "cannot open PDB file " + path);
    std::vector<Record> R;
    std::string line;
    while (std::getline(f, line)) {
        const std::string head = line.substr(0, 6);
        if (head != "ATOM  " && head != "HETATM") continue;
        if (line.size() > 16 && line[16] != ' ' && line[16] != 'A') continue;  // alternate location
        std::string e = line.size() > 76 ? line.substr(76, 4) : "";
        if (e.find_first_not_of(" \t\n\v\f\r") == std::string::npos) e.clear();
        switch (e.size()) {
            case 4:
                if (e[3] == '\r') e[3] = ' ';
                // fall through
            case 3:
                if (e[2] == '\r')
                    e.erase(2);
                else
                    break;
                // fall through
            case 2:
                e.append("  ");
                break;
            case 0:
                e = (line.size() > 13 ? line.substr(12, 2) : std::string("  ")) + "  ";
                if (e[0] >= '0' && e[0] <= '9') e[0] = ' ';
                break;
            default:
                break;
        }
        auto field = [&](size_t a, size_t n) { return line.size() > a ? line.substr(a, n) : std::string(); };
        Record r;
        r.x = atof(field(30, 8).c_str()) / 10.0;
        r.y = atof(field(38, 8).c_str()) / 10.0;
        r.z = atof(field(46, 8).c_str()) / 10.0;
        r.elem4 = e;
        r.name = field(12, 4);
        r.res_name = field(17, 3);
        r.res_no = field(22, 4);
        r.serial = field(6, 5);
        r.chain = line.size() > 21 ? line[21] : '\0';
        r.occ = atof(field(54, 6).c_str());
        r.bfac = atof(field(60, 6).c_str());
        r.seg = field(72, 4);
        R.push_back(r);
    }
    if (R.empty()) throw Error(kErrInvalidArgs, "The file does not contains any valid ATOM or HETATM lines: " + path);
    PdbAtoms out;
    std::vector<std::string> implicit(R.size());
    for (size_t i = 0; i < R.size();) {
        size_t n = R.size() - i;
        bool has_h = false;
        for (size_t j = i; j < R.size(); j++) {
            if (R[j].res_no != R[i].res_no || R[j].res_name != R[i].res_name) {
                n = j - i;
                break;
            }
            if (R[j].elem4 == " H  " || R[j].elem4 == " h  ") has_h = true;
        }
        std::vector<std::string> labels;
        if (!has_h && implicit_labels(R, i, n, labels)) {
            for (size_t k = 0; k < n; k++) implicit[i + k] = labels[k];
            out.implicit_atoms += (int)n;
            out.implicit_residues++;
        }
        i += n;
    }
    for (size_t i = 0; i < R.size(); i++) {
        std::string key = implicit[i].empty() ? R[i].elem4 : implicit[i];
        if (key.size() >= 2) {
            key[0] = (char)tolower((unsigned char)key[0]);
            key[1] = (char)tolower((unsigned char)key[1]);
        }
        const int s = species_index(key);
        if (s < 0)
            throw Error(kErrUnimplemented, "no X-ray form factor table entry for atom type \"" + key + "\" (" + path + ")");
        out.xyz.push_back(R[i].x);
        out.xyz.push_back(R[i].y);
        out.xyz.push_back(R[i].z);
        out.species.push_back(s);
        out.serial.push_back(R[i].serial);
        out.name.push_back(R[i].name);
        out.res_name.push_back(R[i].res_name);
        out.res_no.push_back(R[i].res_no);
        out.seg.push_back(R[i].seg);
        out.elem4.push_back(R[i].elem4);
        out.chain.push_back(R[i].chain);
        out.occ.push_back(R[i].occ);
        out.bfac.push_back(R[i].bfac);
    }
    return out;
}

}  // namespace dplus
