// ampfile.cpp -- amplitude grid files.
//
// Two formats hold a Jacobian grid's 2*P doubles (complex interleaved, node
// order of IndexFromIndices):
//  * legacy ".amp" stream (JacobianSphereGrid::ExportBinaryData, Backend/Backend/Grid.cpp:847-894,
//    reader Grid.cpp:272-345): "#@", u32 header length, "\n", '#'-prefixed
//    header lines, then text "13\n16\n<N>\n3\n", the f64 step and the data;
//  * ".ampj": a zip archive with stored (uncompressed) entries grid.dat,
//    criticalinfo.json {qmax, gridSize, stepSize, ...} and header.json
//    (Amplitude::WriteAmplitudeToFile, Amplitude.cpp:3119-3190, which uses
//    ZipLib's StoreMethod; PythonInterface/dplus/Amplitudes.py:319-337, 705-750).

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "device.hpp"
#include "json.hpp"

namespace dplus {

namespace {

uint32_t crc32_bytes(const char *p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ (uint8_t)p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

std::string slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(kErrFileNotFound, "cannot open " + path);
    std::ostringstream o;
    o << f.rdbuf();
    return o.str();
}

uint16_t rd16(const std::string &z, size_t o) {
    if (o + 2 > z.size()) throw Error(kErrGeneral, "truncated zip archive");
    uint16_t v;
    memcpy(&v, z.data() + o, 2);
    return v;
}

uint32_t rd32(const std::string &z, size_t o) {
    if (o + 4 > z.size()) throw Error(kErrGeneral, "truncated zip archive");
    uint32_t v;
    memcpy(&v, z.data() + o, 4);
    return v;
}

// entries of a zip archive with stored members, read through the central directory
std::vector<std::pair<std::string, std::string>> unzip_stored(const std::string &z, const std::string &path) {
    if (z.size() < 22) throw Error(kErrGeneral, path + " is not a zip archive");
    size_t eocd = std::string::npos;
    for (size_t i = z.size() - 22 + 1; i-- > 0;)
        if (rd32(z, i) == 0x06054b50u) {
            eocd = i;
            break;
        }
    if (eocd == std::string::npos) throw Error(kErrGeneral, path + " is not a zip archive");
    const int n = rd16(z, eocd + 10);
    size_t cd = rd32(z, eocd + 16);
    std::vector<std::pair<std::string, std::string>> out;
    for (int e = 0; e < n; e++) {
        if (rd32(z, cd) != 0x02014b50u) throw Error(kErrGeneral, "bad zip central directory in " + path);
        const uint16_t method = rd16(z, cd + 10);
        const uint32_t crc = rd32(z, cd + 16), csize = rd32(z, cd + 20), usize = rd32(z, cd + 24);
        const uint16_t nl = rd16(z, cd + 28), xl = rd16(z, cd + 30), cl = rd16(z, cd + 32);
        const uint32_t loc = rd32(z, cd + 42);
        if (cd + 46 + nl > z.size()) throw Error(kErrGeneral, "truncated zip archive " + path);
        std::string name = z.substr(cd + 46, nl);
        cd += 46 + nl + xl + cl;
        if (method != 0)
            throw Error(kErrUnimplemented, "zip member " + name + " of " + path + " is compressed; amplitude files store members");
        if (rd32(z, loc) != 0x04034b50u) throw Error(kErrGeneral, "bad zip local header in " + path);
        const size_t data = loc + 30 + rd16(z, loc + 26) + rd16(z, loc + 28);
        if (data + csize > z.size() || csize != usize) throw Error(kErrGeneral, "truncated zip member " + name);
        std::string body = z.substr(data, csize);
        if (crc32_bytes(body.data(), body.size()) != crc) throw Error(kErrGeneral, "CRC mismatch in zip member " + name);
        out.emplace_back(std::move(name), std::move(body));
    }
    return out;
}

void load_into(Grid &g, const char *src, size_t bytes) {
    g.ensure_host();
    const size_t need = (size_t)(2 * g.P) * sizeof(double);
    if (bytes < need) throw Error(kErrGeneral, "amplitude data holds " + std::to_string(bytes) + " bytes, the grid needs " +
                                                   std::to_string(need));
    memcpy(g.host_data.data(), src, need);
    for (size_t i = 0; i < (size_t)(2 * g.P); i++)
        if (!std::isfinite(g.host_data[i]))  // JacobianSphereGrid::Validate (Grid.cpp:987-1005)
            throw Error(kErrGeneral, "amplitude data has a non-finite value at " + std::to_string(i));
}

}  // namespace

std::string amp_legacy_bytes(Grid &g, const std::string &header, hipStream_t s) {
    g.download(s);
    std::string h;
    std::istringstream in(header);
    for (std::string line; std::getline(in, line);) h += (line.empty() || line[0] != '#' ? "# " : "") + line + "\n";
    if (!h.empty()) h.pop_back();
    std::string o = "#@";
    const uint32_t headlen = (uint32_t)(2 + sizeof(uint32_t) + 1 + h.size() + 1);
    o.append((const char *)&headlen, sizeof headlen);
    o += "\n" + h + "\n";
    o += "13\n16\n" + std::to_string(g.N) + "\n3\n";
    o.append((const char *)&g.step, sizeof(double));
    o.append((const char *)g.host_data.data(), g.host_data.size() * sizeof(double));
    return o;
}

void amp_write(Grid &g, const std::string &path, const std::string &title, hipStream_t s) {
    const bool legacy = path.size() >= 4 && path.compare(path.size() - 4, 4, ".amp") == 0;
    std::string z;
    if (legacy) {
        z = amp_legacy_bytes(g, title, s);
    } else {
        g.download(s);
        const std::string info = "{\"qmax\":" + json::num(g.qmax) + ",\"qmin\":0,\"gridSize\":" + std::to_string(g.grid_size) +
                                 ",\"stepSize\":" + json::num(g.step) + ",\"Params\":[]}";
        const std::string head = "{\"Header\":{\"Title\":" + json::quote(title) + ",\"qMax\":" + json::num(g.qmax) +
                                 ",\"StepSize\":" + json::num(g.step) + ",\"N\":" + std::to_string(g.grid_size) +
                                 ",\"Used Grid\":true}}";
        struct Entry {
            std::string name;
            const char *p;
            size_t n;
            uint32_t crc = 0, off = 0;
        };
        std::vector<Entry> es = {{"grid.dat", (const char *)g.host_data.data(), g.host_data.size() * sizeof(double)},
                                 {"criticalinfo.json", info.data(), info.size()},
                                 {"header.json", head.data(), head.size()}};
        auto u16 = [&](uint16_t v) { z.append((const char *)&v, 2); };
        auto u32 = [&](uint32_t v) { z.append((const char *)&v, 4); };
        for (auto &e : es) {
            if (e.n >= 0xFFFFFFFFull) throw Error(kErrUnimplemented, "grid too large for a zip32 archive");
            e.crc = crc32_bytes(e.p, e.n);
            e.off = (uint32_t)z.size();
            u32(0x04034b50u);
            u16(20), u16(0), u16(0), u16(0), u16(0x21);
            u32(e.crc), u32((uint32_t)e.n), u32((uint32_t)e.n);
            u16((uint16_t)e.name.size()), u16(0);
            z += e.name;
            z.append(e.p, e.n);
        }
        const uint32_t cd = (uint32_t)z.size();
        for (auto &e : es) {
            u32(0x02014b50u);
            u16(20), u16(20), u16(0), u16(0), u16(0), u16(0x21);
            u32(e.crc), u32((uint32_t)e.n), u32((uint32_t)e.n);
            u16((uint16_t)e.name.size()), u16(0), u16(0), u16(0), u16(0);
            u32(0), u32(e.off);
            z += e.name;
        }
        const uint32_t cdlen = (uint32_t)z.size() - cd;
        u32(0x06054b50u);
        u16(0), u16(0), u16((uint16_t)es.size()), u16((uint16_t)es.size());
        u32(cdlen), u32(cd), u16(0);
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error(kErrFileNotFound, "cannot write " + path);
    f.write(z.data(), (std::streamsize)z.size());
    if (!f) throw Error(kErrGeneral, "write failed: " + path);
}

// AmpGridAmplitude::ReadAmplitudeFromFile (Amplitude.cpp:3724-3796): the grid
// of a .amp/.ampj file, uploaded and splined, at the requested precision
Grid *amp_read(const std::string &path, int precision, std::string &header, hipStream_t s) {
    const std::string z = slurp(path);
    const bool legacy = path.size() >= 4 && path.compare(path.size() - 4, 4, ".amp") == 0;
    std::unique_ptr<Grid> g;
    if (legacy) {
        size_t pos = 0;
        uint32_t offset = 0;
        if (z.size() >= 2 && z[0] == '#') {
            if (z[1] == '@') {
                if (z.size() < 6) throw Error(kErrGeneral, "truncated amplitude file " + path);
                memcpy(&offset, z.data() + 2, 4);
                pos = z.find('\n', 6);
                pos = pos == std::string::npos ? z.size() : pos + 1;
            } else if (z[1] != '\n') {
                size_t e = z.find('\n', 2);
                header += z.substr(2, e == std::string::npos ? std::string::npos : e - 2) + "\n";
                pos = e == std::string::npos ? z.size() : e + 1;
            }
        }
        while (pos < z.size() && z[pos] == '#') {
            size_t e = z.find('\n', pos);
            if (e == std::string::npos) e = z.size();
            header += z.substr(pos, e - pos) + "\n";
            pos = std::min(z.size(), e + 1);
        }
        if (offset > 0) pos = offset;
        // version, bytes per element, actual size, extras, then the step after the newline
        std::istringstream in(z.substr(pos, 200));
        unsigned version = 0, nbytes = 0, actual = 0, extras = 0;
        in >> version >> nbytes >> actual >> extras;
        if (!in || version < 13 || nbytes != 16 || extras != 3 || actual < 4)
            throw Error(kErrGeneral, "unsupported amplitude file " + path + " (version " + std::to_string(version) + ")");
        const size_t nl = z.find('\n', pos + (size_t)in.tellg());
        if (nl == std::string::npos || nl + 1 + sizeof(double) > z.size()) throw Error(kErrGeneral, "truncated amplitude file " + path);
        double step;
        memcpy(&step, z.data() + nl + 1, sizeof step);
        const int gs = (int)(actual - extras) * 2;
        if (!(step > 0)) throw Error(kErrGeneral, "Amp file is not valid");
        g.reset(new Grid(gs, step * gs / 2.0, precision));
        load_into(*g, z.data() + nl + 1 + sizeof(double), z.size() - (nl + 1 + sizeof(double)));
    } else {
        auto es = unzip_stored(z, path);
        const std::string *dat = nullptr, *info = nullptr;
        for (auto &e : es) {
            if (e.first == "grid.dat") dat = &e.second;
            if (e.first == "criticalinfo.json") info = &e.second;
            if (e.first == "header.json") header = e.second;
        }
        if (!dat || !info) throw Error(kErrFileNotFound, path + " lacks grid.dat or criticalinfo.json");
        json::Value v = json::parse(*info);
        const double qmax = v.at("qmax").number();
        const int gs = (int)v.at("gridSize").number();
        g.reset(new Grid(gs, qmax, precision));
        load_into(*g, dat->data(), dat->size());
    }
    g->upload(s);
    return g.release();
}

}  // namespace dplus

using namespace dplus;

extern "C" {

int dplus_amp_read(const char *path, int precision, dplus_grid **out, char **header) {
    try {
        if (!path || !out) throw Error(kErrInvalidArgs, "null argument");
        Device::get().require();
        std::string h;
        Grid *g = amp_read(path, precision, h, Device::get().stream);
        check_hip(hipStreamSynchronize(Device::get().stream), "amp read");
        *out = (dplus_grid *)g;
        if (header) {
            *header = (char *)malloc(h.size() + 1);
            if (*header) memcpy(*header, h.c_str(), h.size() + 1);
        }
        return kOk;
    } catch (const Error &e) {
        return set_error(e.code, e.what());
    } catch (const std::exception &e) {
        return set_error(kErrGeneral, e.what());
    }
}

int dplus_amp_write(dplus_grid *g, const char *path, const char *header) {
    try {
        if (!g || !path) throw Error(kErrInvalidArgs, "null argument");
        Device::get().require();
        amp_write(*(Grid *)g, path, header ? header : "", Device::get().stream);
        return kOk;
    } catch (const Error &e) {
        return set_error(e.code, e.what());
    } catch (const std::exception &e) {
        return set_error(kErrGeneral, e.what());
    }
}

}  // extern "C"
