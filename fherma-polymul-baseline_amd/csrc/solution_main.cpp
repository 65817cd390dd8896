// `solution <point directory>` -- the point-directory harness of the
// challenge, written against fherma.h.  Wire contract (reference main.cpp):
//   manifest.json            keys "N", "W", "L", "cases"          (read)
//   point/q.bin              L little-endian u32                  (read)
//   cases/NNNNNN/{a,b}.bin   N*L little-endian u32 each           (read)
//   out/NNNNNN/c.bin         N*L little-endian u32                (written)
//   out/results.json         {"init_s":s,"cases":[{"i":..,"seconds":..,"status":"ok"|"crashed"...}]}
// Only fherma_run is inside the timed window; results.json is rewritten after
// every case so a killed run keeps the cases it finished.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>
#include <vector>

#include "fherma.h"

namespace fs = std::filesystem;
using Clock = std::chrono::steady_clock;

static std::string read_file(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + path.string());
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

// Value of "key": <number> in our own two-level manifest (no JSON library).
static double manifest_number(const std::string& text, const char* key) {
    const std::string needle = std::string("\"") + key + "\"";
    const size_t at = text.find(needle);
    if (at == std::string::npos) throw std::runtime_error(std::string("manifest.json lacks ") + key);
    const size_t colon = text.find(':', at + needle.size());
    if (colon == std::string::npos) throw std::runtime_error(std::string("manifest.json: bad ") + key);
    return std::strtod(text.c_str() + colon + 1, nullptr);
}

static fherma::Tensor<uint32_t> read_words(const fs::path& path, std::vector<int64_t> shape) {
    fherma::Tensor<uint32_t> t;
    t.shape = std::move(shape);
    const std::string bytes = read_file(path);
    const size_t count = size_t(t.count());
    if (bytes.size() != count * sizeof(uint32_t))
        throw std::runtime_error(path.filename().string() + ": " + std::to_string(bytes.size()) +
                                 " bytes, expected " + std::to_string(count * sizeof(uint32_t)));
    t.data.resize(count);
    std::memcpy(t.data.data(), bytes.data(), bytes.size());  // little-endian host
    return t;
}

static std::string case_name(size_t i) {
    char name[32];
    std::snprintf(name, sizeof(name), "%06zu", i);
    return name;
}

static std::string json_escape(const std::string& s) {
    std::string out;
    for (char ch : s.substr(0, 200)) {
        if (ch == '"' || ch == '\\') out += '\\';
        out += ch == '\n' ? ' ' : ch;
    }
    return out;
}

static void write_report(const fs::path& out_dir, double init_s, const std::vector<std::string>& entries) {
    std::ostringstream doc;
    doc.setf(std::ios::fixed);
    doc.precision(9);
    doc << "{\"init_s\":" << init_s << ",\"cases\":[";
    for (size_t i = 0; i < entries.size(); ++i) doc << (i ? "," : "") << entries[i];
    doc << "]}";
    std::ofstream(out_dir / "results.json") << doc.str();
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: solution <point directory>\n";
        return 2;
    }
    const fs::path root = argv[1];
    const std::string manifest = read_file(root / "manifest.json");
    fherma::Point point;
    point.N = uint32_t(manifest_number(manifest, "N"));
    point.W = uint32_t(manifest_number(manifest, "W"));
    point.L = uint32_t(manifest_number(manifest, "L"));
    point.q = read_words(root / "point" / "q.bin", {int64_t(point.L)});
    const size_t cases = size_t(manifest_number(manifest, "cases"));
    const fs::path out_dir = root / "out";
    fs::create_directories(out_dir);

    const auto t0 = Clock::now();
    void* state = fherma_init(point);  // init failures end the process, as in the reference harness
    const double init_s = std::chrono::duration<double>(Clock::now() - t0).count();
    std::vector<std::string> entries;
    write_report(out_dir, init_s, entries);

    for (size_t i = 0; i < cases; ++i) {
        const fs::path in_dir = root / "cases" / case_name(i);
        std::ostringstream entry;
        entry.setf(std::ios::fixed);
        entry.precision(9);
        try {
            fherma::Inputs in;
            try {
                in.a = read_words(in_dir / "a.bin", {int64_t(point.N), int64_t(point.L)});
                in.b = read_words(in_dir / "b.bin", {int64_t(point.N), int64_t(point.L)});
            } catch (const std::exception& e) {
                throw std::runtime_error(std::string("reading the case: ") + e.what());
            }
            const auto start = Clock::now();
            fherma::Outputs result = fherma_run(state, in);
            const double seconds = std::chrono::duration<double>(Clock::now() - start).count();
            try {
                const fs::path answer_dir = out_dir / case_name(i);
                fs::create_directories(answer_dir);
                std::ofstream(answer_dir / "c.bin", std::ios::binary)
                    .write(reinterpret_cast<const char*>(result.c.data.data()),
                           std::streamsize(result.c.data.size() * sizeof(uint32_t)));
            } catch (const std::exception& e) {
                throw std::runtime_error(std::string("writing the answer: ") + e.what());
            }
            entry << "{\"i\":" << i << ",\"seconds\":" << seconds << ",\"status\":\"ok\"}";
        } catch (const std::exception& e) {
            entry << "{\"i\":" << i << ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" << json_escape(e.what())
                  << "\"}";
        }
        entries.push_back(entry.str());
        write_report(out_dir, init_s, entries);
    }
    fherma_free(state);
    write_report(out_dir, init_s, entries);
    return 0;
}
