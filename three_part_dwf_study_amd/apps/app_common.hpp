// app_common.hpp -- host support of the generate/analyze programs: the parameters.txt
// reader, NPY v1 <f8 and WILCFG1 ensemble files, small helpers.  Behaviour and
// messages follow the reference (common/config.hpp:11-109,
// domain_wall/include/npy_io.hpp:17-158, wilson/include/config_io.hpp:13-51) so the
// programs are drop-ins for its executables.
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace app {

// "key = value" lines, '#' comments; strict numeric getters (the whole token must parse).
class Parameters {
public:
    explicit Parameters(const std::string& path = "parameters.txt")
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Cannot open parameter file: " + path);
        std::string line;
        int lineno = 0;
        while (std::getline(in, line)) {
            ++lineno;
            const auto hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            line = trim(line);
            if (line.empty()) continue;
            const auto eq = line.find('=');
            if (eq == std::string::npos)
                throw std::runtime_error(path + ":" + std::to_string(lineno) + ": expected key = value");
            const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            if (k.empty() || v.empty())
                throw std::runtime_error(path + ":" + std::to_string(lineno) + ": empty key or value");
            kv_[k] = v;
        }
    }
    bool has(const std::string& k) const { return kv_.count(k) != 0; }
    std::string text(const std::string& k) const { return need(k); }
    int integer(const std::string& k) const { return parse<int>(k, [](const std::string& s, size_t* u) { return std::stoi(s, u); }); }
    unsigned long long unsigned_integer(const std::string& k) const
    {
        return parse<unsigned long long>(k, [](const std::string& s, size_t* u) { return std::stoull(s, u); });
    }
    double real(const std::string& k) const { return parse<double>(k, [](const std::string& s, size_t* u) { return std::stod(s, u); }); }
    // Optional keys this build adds (gpu.*): the default when absent.
    std::string text_or(const std::string& k, const std::string& d) const { return has(k) ? text(k) : d; }

private:
    std::unordered_map<std::string, std::string> kv_;
    static std::string trim(const std::string& s)
    {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) ++a;
        while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
        return s.substr(a, b - a);
    }
    std::string need(const std::string& k) const
    {
        auto it = kv_.find(k);
        if (it == kv_.end()) throw std::runtime_error("Missing required parameter: " + k);
        return it->second;
    }
    template <class T, class F> T parse(const std::string& k, F f) const
    {
        const std::string v = need(k);
        size_t used = 0;
        T r{};
        try {
            r = f(v, &used);
        } catch (const std::exception&) {
            throw std::runtime_error("Invalid numeric value for " + k + ": " + v);
        }
        if (used != v.size()) throw std::runtime_error("Invalid numeric value for " + k + ": " + v);
        return r;
    }
};

// ---------------------------------------------------------------- NPY v1, <f8
inline void save_npy(const std::string& path, const std::vector<size_t>& shape, const std::vector<double>& data)
{
    size_t n = 1;
    for (size_t d : shape) n *= d;
    if (n != data.size()) throw std::runtime_error("NPY shape does not match data size");
    std::string sh = "(";
    for (size_t i = 0; i < shape.size(); ++i) {
        sh += std::to_string(shape[i]);
        if (shape.size() == 1 || i + 1 < shape.size()) sh += ",";
        if (i + 1 < shape.size()) sh += " ";
    }
    sh += ")";
    std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': " + sh + ", }";
    size_t pad = 16 - (10 + header.size() + 1) % 16;  // magic(6)+version(2)+len(2)
    if (pad == 16) pad = 0;
    header.append(pad, ' ');
    header.push_back('\n');
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open " + path + " for writing");
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    out.write(magic, 8);
    const uint16_t hl = (uint16_t)header.size();
    const char len[2] = {(char)(hl & 0xff), (char)(hl >> 8)};
    out.write(len, 2);
    out.write(header.data(), (std::streamsize)header.size());
    out.write(reinterpret_cast<const char*>(data.data()), (std::streamsize)(data.size() * sizeof(double)));
    if (!out) throw std::runtime_error("Failed to write " + path);
}

struct NpyArray {
    std::vector<size_t> shape;
    std::vector<double> data;
};

inline NpyArray load_npy(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open " + path);
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error("Not an NPY file: " + path);
    size_t hl = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        in.read(reinterpret_cast<char*>(b), 2);
        hl = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        in.read(reinterpret_cast<char*>(b), 4);
        hl = b[0] | (b[1] << 8) | (b[2] << 16) | ((size_t)b[3] << 24);
    }
    std::string header(hl, '\0');
    in.read(&header[0], (std::streamsize)hl);
    if (!in) throw std::runtime_error("Truncated NPY header");
    if (header.find("'<f8'") == std::string::npos) throw std::runtime_error("NPY dtype must be <f8");
    if (header.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error("Fortran-ordered NPY arrays are not supported");
    const auto a = header.find('('), b = header.find(')', a);
    if (a == std::string::npos || b == std::string::npos) throw std::runtime_error("NPY shape missing");
    NpyArray arr;
    std::stringstream ss(header.substr(a + 1, b - a - 1));
    std::string tok;
    size_t n = 1;
    while (std::getline(ss, tok, ',')) {
        tok.erase(std::remove_if(tok.begin(), tok.end(), [](unsigned char c) { return std::isspace(c); }), tok.end());
        if (tok.empty()) continue;
        arr.shape.push_back(std::stoull(tok));
        n *= arr.shape.back();
    }
    arr.data.resize(n);
    in.read(reinterpret_cast<char*>(arr.data.data()), (std::streamsize)(n * sizeof(double)));
    if (!in) throw std::runtime_error("Truncated NPY data");
    return arr;
}

// ------------------------------------------------------------- WILCFG1 ensembles
// 40-byte header {char magic[8]; int32 nt, nx; double beta, m0; uint64 count} then
// count x 2 Nt Nx doubles (wilson/include/config_io.hpp:13-51).
struct WilHeader {
    char magic[8];
    int32_t nt, nx;
    double beta, m0;
    uint64_t count;
};
static_assert(sizeof(WilHeader) == 40, "WILCFG1 header is 40 bytes");

inline void save_wilcfg(const std::string& path, int Nt, int Nx, double beta, double m0,
                        const std::vector<std::vector<double>>& configs)
{
    std::filesystem::create_directories(std::filesystem::path(path).parent_path());
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot write " + path);
    WilHeader h{};
    std::memcpy(h.magic, "WILCFG1", 8);
    h.nt = Nt;
    h.nx = Nx;
    h.beta = beta;
    h.m0 = m0;
    h.count = configs.size();
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    for (const auto& c : configs) out.write(reinterpret_cast<const char*>(c.data()), (std::streamsize)(c.size() * 8));
}

inline std::vector<std::vector<double>> load_wilcfg(const std::string& path, int Nt, int Nx)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open " + path);
    WilHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (std::string(h.magic, strnlen(h.magic, 8)) != "WILCFG1") throw std::runtime_error("Wrong ensemble file format");
    if (h.nt != Nt || h.nx != Nx) throw std::runtime_error("Ensemble dimensions do not match");
    std::vector<std::vector<double>> cfgs(h.count, std::vector<double>((size_t)2 * Nt * Nx));
    for (auto& c : cfgs) in.read(reinterpret_cast<char*>(c.data()), (std::streamsize)(c.size() * 8));
    if (!in) throw std::runtime_error("Truncated ensemble file");
    return cfgs;
}

// Solver choice for the GPU (an additive key; the reference's keys are unchanged).
// Learned-preconditioner weights on disk: an NPY <f8 vector of n values (the library's
// float weights, widened).
inline std::vector<float> load_network_weights(const std::string& path, size_t n)
{
    const NpyArray a = load_npy(path);
    if (a.shape.size() != 1 || a.shape[0] != n)
        throw std::runtime_error(path + " must hold " + std::to_string(n) + " network weights");
    return std::vector<float>(a.data.begin(), a.data.end());
}

inline void save_network_weights(const std::string& path, const std::vector<float>& w)
{
    save_npy(path, {w.size()}, std::vector<double>(w.begin(), w.end()));
}

inline int solver_method(const Parameters& p, const std::string& key, int dflt)
{
    const std::string v = p.text_or(key, "");
    if (v.empty()) return dflt;
    if (v == "cg64") return 0;
    if (v == "mixed") return 1;
    if (v == "eo_mixed") return 2;
    if (v == "learned") return 3;
    throw std::runtime_error("Invalid value for " + key + ": " + v + " (cg64, mixed, eo_mixed or learned)");
}

// The HMC force solves: cg64 (the reference's CG), mixed, eo_mixed (two even-odd Dirac
// solves per correction pass, certified on D D^dag) or learned.
inline int force_solver_method(const Parameters& p, const std::string& key) { return solver_method(p, key, 0); }

}  // namespace app
