// cli.cpp -- command-line driver with the flag surface of lightweight_dcx's cli
// (lightweight_dcx/src/executables/cli.cpp:149-305, main at :747-820), running the MI355X engine through the
// C ABI of include/dcx.h.  Flags that only configure the MPI sorters / bucket rounds are accepted and echoed
// in the JSON config: the suffix array is unique, so they cannot change the output.
//
//   dcx_cli --input <file|random> [--textsize N] [--alphabet_size S] [--seed S] [--dcx dc39]
//           [--use-char-packing-samples] [--use-char-packing-merging] [--discarding-threshold 0.7]
//           [--check] [--output sa.bin] [--json-output-path prefix] [--sa-width 5] ...
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "dcx.h"

extern "C" int dcx_gen_text(int kind, uint64_t n, uint64_t seed, uint8_t* out);

namespace {

struct Params {
    std::string input_path;
    uint64_t textsize = 0;
    uint64_t alphabet_size = 2;
    std::string output_path;
    std::string json_output_path;
    uint64_t seed = 0;
    bool check = false;
    std::string dcx_variant = "dc39";
    double discarding_threshold = 0.7;
    bool pack_samples = false, pack_merging = false;
    uint64_t pack_extra_words = 0;
    uint32_t sa_width = 5;  // uint40, cli.cpp:107
    int device = 0;
    std::map<std::string, std::string> echoed;  // accepted MPI-only options
};

// options with a value that only steer the distributed implementation
const char* kValueOnly[] = {"--iteration", "--num-samples-splitters", "--buckets-sample-phase", "--buckets-merging-phase",
                            "--buckets-phase3", "--samples-buckets-phase3", "--avg-chunks-pe", "--atomic-sorter",
                            "--ams-levels", "--splitter-sampling", "--splitter-sorting", "--string-sorter",
                            "--memory_seq_string_sorter", "--lcp-compression-threshold", "--inital-prefix-length"};
const char* kFlagOnly[] = {"--use-random-sampling-splitters", "--balance-blocks-space-efficient-sort",
                           "--use-randomized-chunks", "--rearrange-buckets-balanced", "--use-robust-tie-break",
                           "--use-compressed-buckets", "--use-string-sort", "--use-string-sort-tie-breaking-phase1",
                           "--use-string-sort-tie-breaking-phase4", "--use-loser-tree", "--use-rquick-for-splitters",
                           "--use-binary-search-for-splitters", "--use-lcp-compression", "--use-prefix-doubling"};

int usage(const char* msg) {
    std::cerr << msg << "\nusage: dcx_cli --input <file|random> [--textsize N] [--dcx dcX] [--check] [--output f]"
              << std::endl;
    return 1;
}

bool parse(int argc, char** argv, Params& p, std::string& err) {
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string v;
        auto eq = a.find('=');
        bool has_eq = false;
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_eq = true;
        }
        auto need = [&](std::string& out) -> bool {
            if (has_eq) {
                out = v;
                return true;
            }
            if (i + 1 >= argc) {
                err = "missing value for " + a;
                return false;
            }
            out = argv[++i];
            return true;
        };
        std::string val;
        if (a == "--input") {
            if (!need(p.input_path)) return false;
        } else if (a == "--textsize") {
            if (!need(val)) return false;
            p.textsize = std::stoull(val);
        } else if (a == "--alphabet_size") {
            if (!need(val)) return false;
            p.alphabet_size = std::stoull(val);
        } else if (a == "--output") {
            if (!need(p.output_path)) return false;
        } else if (a == "--json-output-path") {
            if (!need(p.json_output_path)) return false;
        } else if (a == "--seed") {
            if (!need(val)) return false;
            p.seed = std::stoull(val);
        } else if (a == "--check") {
            p.check = true;
        } else if (a == "--dcx") {
            if (!need(p.dcx_variant)) return false;
        } else if (a == "--discarding-threshold") {
            if (!need(val)) return false;
            p.discarding_threshold = std::stod(val);
        } else if (a == "--use-char-packing-samples") {
            p.pack_samples = true;
        } else if (a == "--use-char-packing-merging") {
            p.pack_merging = true;
        } else if (a == "--pack-extra-words") {
            if (!need(val)) return false;
            p.pack_extra_words = std::stoull(val);
        } else if (a == "--sa-width") {
            if (!need(val)) return false;
            p.sa_width = (uint32_t)std::stoul(val);
        } else if (a == "--device") {
            if (!need(val)) return false;
            p.device = std::stoi(val);
        } else {
            bool known = false;
            for (const char* f : kValueOnly)
                if (a == f) {
                    if (!need(val)) return false;
                    p.echoed[a.substr(2)] = val;
                    known = true;
                }
            for (const char* f : kFlagOnly)
                if (a == f) {
                    p.echoed[a.substr(2)] = "true";
                    known = true;
                }
            if (!known) {
                err = "unknown option " + a;
                return false;
            }
        }
    }
    return true;
}

uint32_t parse_x(const std::string& s) {
    if (s.size() < 3 || (s[0] != 'd' && s[0] != 'D') || (s[1] != 'c' && s[1] != 'C')) return 0;
    return (uint32_t)std::strtoul(s.c_str() + 2, nullptr, 10);
}

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    Params p;
    std::string err;
    if (!parse(argc, argv, p, err)) return usage(err.c_str());
    if (p.input_path.empty()) return usage("--input is required");
    if (p.pack_extra_words != 0) return usage("--pack-extra-words supports only 0");
    const uint32_t X = parse_x(p.dcx_variant);

    // read_input (cli.cpp:393-415)
    std::vector<uint8_t> text;
    auto t_start = std::chrono::steady_clock::now();
    if (p.input_path == "random") {
        if (p.alphabet_size < 1 || p.alphabet_size > 256) return usage("--alphabet_size must be in [1, 256]");
        text.resize(p.textsize);
        dcx_gen_text(4, p.textsize, (p.seed & 0xFFFFFFFFull) | ((p.alphabet_size & 0xFF) << 32), text.data());
    } else {
        std::ifstream f(p.input_path, std::ios::binary);
        if (!f) {
            std::cerr << "File " << p.input_path << " does not exist!" << std::endl;
            return 1;
        }
        f.seekg(0, std::ios::end);
        uint64_t size = (uint64_t)f.tellg();
        f.seekg(0);
        if (p.textsize && p.textsize < size) size = p.textsize;
        text.resize(size);
        f.read(reinterpret_cast<char*>(text.data()), (std::streamsize)size);
        if (!f) {
            std::cerr << "error reading " << p.input_path << std::endl;
            return 1;
        }
    }
    double io_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();

    dcx_handle* h = nullptr;
    int rc = dcx_create(&h, p.device, 0);
    if (rc != DCX_OK) {
        std::cerr << "dcx_create: " << dcx_strerror(rc) << std::endl;
        return 1;
    }
    dcx_options o;
    dcx_default_options(&o);
    o.X = X;
    o.pack_samples = p.pack_samples;
    o.pack_merging = p.pack_merging;
    o.discarding_threshold = p.discarding_threshold;
    o.seed = p.seed;
    const uint64_t n = text.size();
    std::vector<uint8_t> sa(n * p.sa_width + 1);
    dcx_stats st;
    auto t0 = std::chrono::steady_clock::now();
    rc = dcx_build_sa(h, text.data(), n, &o, sa.data(), p.sa_width, &st);
    double total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc != DCX_OK) {
        std::cerr << dcx_strerror(rc) << ": " << dcx_last_error(h) << std::endl;
        dcx_destroy(h);
        return 1;
    }
    std::cout << "input_alphabet_size=" << st.sigma << "\n";
    if (st.levels && st.level[0].bits_per_char) {
        std::cout << "packed_chars=" << st.level[0].key_chars << "\n";
        std::cout << "bits_per_char=" << st.level[0].bits_per_char << "\n";
    }
    for (uint32_t l = 0; l < st.levels; l++)
        std::cout << "level=" << l << " n=" << st.level[l].n << " samples=" << st.level[l].samples
                  << " names=" << st.level[l].names << " discarding=" << st.level[l].discarding << "\n";
    std::cout << "total_time=" << total_ms << " ms, device_time=" << st.ms_device << " ms\n";

    int ok = -1;
    double check_ms = 0;
    if (p.check) {
        auto c0 = std::chrono::steady_clock::now();
        rc = dcx_check_sa(h, text.data(), n, sa.data(), p.sa_width, 0, &ok);
        check_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
        if (rc != DCX_OK) {
            std::cerr << dcx_strerror(rc) << std::endl;
            dcx_destroy(h);
            return 1;
        }
        std::cout << (ok ? "Correct SA!" : "ERROR: Not a correct SA!") << "\n";
        std::cout << "SA_ok=" << ok << std::endl;
    }
    if (!p.output_path.empty()) {
        std::cout << "Writing the SA to " << p.output_path << "\n";
        std::ofstream f(p.output_path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(sa.data()), (std::streamsize)(n * p.sa_width));
        if (!f) {
            std::cerr << "error writing " << p.output_path << std::endl;
            dcx_destroy(h);
            return 1;
        }
        std::cout << "Finished writing the SA \n";
    }
    if (!p.json_output_path.empty()) {
        std::ostringstream cfg;
        cfg << "{\"input\":\"" << json_escape(p.input_path) << "\",\"textsize\":" << n << ",\"dcx\":\""
            << json_escape(p.dcx_variant) << "\",\"discarding_threshold\":" << p.discarding_threshold
            << ",\"use_char_packing_samples\":" << (p.pack_samples ? "true" : "false")
            << ",\"use_char_packing_merging\":" << (p.pack_merging ? "true" : "false") << ",\"seed\":" << p.seed
            << ",\"num_pe\":1";
        for (auto& kv : p.echoed) cfg << ",\"" << json_escape(kv.first) << "\":\"" << json_escape(kv.second) << "\"";
        cfg << "}";
        std::ofstream ft(p.json_output_path + "_timer.json");
        ft << "[{\"config\":" << cfg.str() << ",\"timer\":{\"io\":" << io_ms << ",\"total_time\":" << total_ms
           << ",\"device_time\":" << st.ms_device << ",\"phase_00_preparation\":" << st.ms_phase[0]
           << ",\"phase_01_02_samples_ranks\":" << st.ms_phase[1] << ",\"phase_03_recursion\":" << st.ms_phase[2]
           << ",\"phase_04_merge\":" << st.ms_phase[3] << ",\"output\":" << st.ms_phase[4]
           << ",\"check_SA\":" << check_ms << "}}]\n";
        std::ofstream fc(p.json_output_path + "_counter.json");
        fc << "[{\"config\":" << cfg.str() << ",\"counter\":{\"input_alphabet_size\":" << st.sigma
           << ",\"levels\":" << st.levels << ",\"alg_bytes\":" << st.alg_bytes
           << ",\"peak_device_bytes\":" << st.peak_device_bytes << ",\"SA_ok\":" << ok << ",\"string_sizes\":[";
        for (uint32_t l = 0; l < st.levels; l++) fc << (l ? "," : "") << st.level[l].n;
        fc << "],\"highest_ranks\":[";
        for (uint32_t l = 0; l < st.levels; l++) fc << (l ? "," : "") << st.level[l].names;
        fc << "],\"use_discarding\":[";
        for (uint32_t l = 0; l < st.levels; l++) fc << (l ? "," : "") << st.level[l].discarding;
        fc << "]}}]\n";
    }
    dcx_destroy(h);
    return (p.check && !ok) ? 1 : 0;
}
