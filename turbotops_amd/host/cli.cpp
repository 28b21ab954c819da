// SPDX-License-Identifier: MIT
// turbotops — the `run` command of the reference's command-line tool
// (src/turbotops.cpp:784-933) over the MI355X front end: the same options,
// input formats, diagram text and exit codes, so that scripts written for
// the reference tool run unchanged and print the same diagram bytes.
//
//   turbotops run <input> [options]
//
// Inputs (src/turbotops.cpp:126-203):
//   text    one point per line; ',', tab and ';' separate like spaces; '#'
//           starts a comment; blank lines are skipped; a UTF-8 BOM is skipped
//   .bin    raw little-endian float32, row-major (needs --dim)
//   --edges weighted edge list 'i j w' (0-based; ',' and tab separate);
//           self loops are dropped, a repeated edge keeps its smallest weight
// Output (src/turbotops.cpp:205-230): one bar per line, 'dim birth death',
// numbers as %.9g, 'inf' for essential deaths, sorted by (dim, birth, death).
// Errors print 'turbotops: <message>' on stderr and exit with status 2.
//
// `gen` and `test` (point generators and the reference's self-test) are not
// part of this tool.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "turbotops_gpu.hpp"

using namespace tops;

namespace {

const char* kHelp =
    "usage:\n"
    "  turbotops run <input> [options]   compute a persistence diagram\n"
    "  turbotops --version               print the version and exit\n"
    "\n"
    "input:\n"
    "  text     one point per line; ',', tab, ';' or spaces between coordinates;\n"
    "           '#' starts a comment\n"
    "  .bin     raw little-endian float32, row-major; needs --dim\n"
    "  --edges  the input is a weighted edge list 'i j w' (0-based), one per line\n"
    "\n"
    "options:\n"
    "  --maxdim D          highest homology dimension (default 1)\n"
    "  --threshold T       filtration cutoff, inclusive (default inf)\n"
    "  --metric M          l1 | l2 | linf | lp:<p>, p >= 1 (default l2)\n"
    "  --graph full|knn    all pairs, or the exact kNN graph (default full)\n"
    "  -k N                neighbours for --graph knn (default 15)\n"
    "  --dim D             coordinates per point for .bin input\n"
    "  --mode M            auto | boundary | coboundary (default auto)\n"
    "  --coboundary        same as --mode coboundary\n"
    "  --boundary, --no-coboundary   same as --mode boundary\n"
    "  --engine E          auto | reducer (= coboundary) | redzed (= boundary)\n"
    "  --threads N         host threads for the reduction (default: all cores)\n"
    "  --memory-budget B   bytes for the host matrix, k/m/g suffix allowed\n"
    "                      (default: half of RAM)\n"
    "  --output FILE       write the diagram to FILE instead of stdout\n"
    "  --quiet             no summary on stderr\n"
    "  --approximate, --verify-recall S, --hnsw-m N, --ef-construction N,\n"
    "  --ef-search N       HNSW kNN: accepted, but HNSW is not in this build\n"
    "\n"
    "output: one bar per line, 'dim birth death' ('inf' for essential classes).\n"
    "The distances, radius graphs, kNN lists, H0 and the edge orders run on the GPU.\n";

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const std::string& m) { throw Fail(m); }

// strtoull/strtof with the whole token consumed (src/turbotops.cpp:102-123):
// the same C library conversions, so a value reads identically in both tools.
u64 to_u64(const std::string& s, const char* what) {
    if (s.empty()) fail(std::string("bad ") + what + " '" + s + "'");
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
    if (errno != 0 || end == nullptr || *end != '\0') fail(std::string("bad ") + what + " '" + s + "'");
    return u64(v);
}

u32 to_u32(const std::string& s, const char* what) {
    const u64 v = to_u64(s, what);
    if (v > 0xFFFFFFFFull) fail(std::string("bad ") + what + " '" + s + "' (out of range)");
    return u32(v);
}

f32 to_f32(const std::string& s, const char* what) {
    if (s.empty()) fail(std::string("bad ") + what + " '" + s + "'");
    char* end = nullptr;
    const f32 v = std::strtof(s.c_str(), &end);
    if (end == nullptr || *end != '\0') fail(std::string("bad ") + what + " '" + s + "'");
    return v;
}

// A byte count with an optional k/m/g (binary) suffix (src/turbotops.cpp:84-100).
u64 to_bytes(const std::string& s) {
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    double scale = 1.0;
    if (end != nullptr && *end != '\0') {
        const char c = char(*end | 0x20);
        if (c == 'k') scale = 1024.0;
        else if (c == 'm') scale = 1048576.0;
        else if (c == 'g') scale = 1073741824.0;
        else fail("bad byte count '" + s + "'");
    }
    if (!(v > 0)) fail("bad byte count '" + s + "'");
    return u64(v * scale);
}

bool has_suffix(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

// Splits a line into whitespace-separated tokens after mapping each byte of
// `seps` to a space and cutting the line at '#'.
std::vector<std::string> tokens(std::string line, const char* seps) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        const bool sep = std::isspace(static_cast<unsigned char>(ch)) || std::char_traits<char>::find(seps, std::char_traits<char>::length(seps), ch);
        if (sep) {
            if (!cur.empty()) out.push_back(std::move(cur)), cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(std::move(cur));
    return out;
}

// Text points (src/turbotops.cpp:126-155).
std::vector<f32> load_text(const std::string& path, u32& n, u32& dim) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("cannot open " + path);
    std::vector<f32> x;
    std::string line;
    n = dim = 0;
    bool first = true;
    while (std::getline(in, line)) {
        if (first && line.size() >= 3 && (unsigned char)line[0] == 0xEF && (unsigned char)line[1] == 0xBB &&
            (unsigned char)line[2] == 0xBF)
            line.erase(0, 3);
        first = false;
        const auto tok = tokens(line, ",\t;");
        if (tok.empty()) continue;
        for (const auto& t : tok) x.push_back(to_f32(t, "coordinate"));
        if (dim == 0) dim = u32(tok.size());
        else if (tok.size() != dim) fail(path + ": inconsistent row length at point " + std::to_string(n));
        ++n;
    }
    return x;
}

// .bin points (src/turbotops.cpp:157-169).
std::vector<f32> load_bin(const std::string& path, u32 dim, u32& n) {
    if (dim == 0) fail("--dim is required for .bin input");
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) fail("cannot open " + path);
    const u64 bytes = u64(in.tellg());
    const u64 row = u64(dim) * sizeof(f32);
    if (bytes % row != 0) fail(path + ": size is not a multiple of 4 * dim");
    n = u32(bytes / row);
    std::vector<f32> x(bytes / sizeof(f32));
    in.seekg(0);
    in.read(reinterpret_cast<char*>(x.data()), std::streamsize(bytes));
    return x;
}

// Edge list (src/turbotops.cpp:171-203).
CsrGraph load_edges(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("cannot open " + path);
    std::vector<std::vector<std::pair<u32, f32>>> rows;
    std::string line;
    u64 lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const auto tok = tokens(line, ",\t");
        if (tok.empty()) continue;
        const std::string where = path + ":" + std::to_string(lineno);
        if (tok.size() < 3) fail(where + ": expected 'i j w'");
        const u64 i = to_u64(tok[0], "vertex"), j = to_u64(tok[1], "vertex");
        const f32 w = to_f32(tok[2], "weight");
        if (i == j) continue;
        if (!(w >= 0) || !std::isfinite(w)) fail(where + ": bad weight");
        const u64 top = std::max(i, j);
        if (top >= 0xFFFFFFFFull) fail("vertex id too large");
        if (rows.size() <= top) rows.resize(top + 1);
        rows[i].emplace_back(u32(j), w);
        rows[j].emplace_back(u32(i), w);
    }
    for (auto& r : rows) {
        // by neighbour, then weight: the first of each neighbour is its smallest weight
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first == b.first; }),
                r.end());
    }
    return CsrGraph::from_rows(rows);
}

// The canonical diagram text (src/turbotops.cpp:205-230).
void print_diagram(FILE* f, const Barcode& bc) {
    struct Bar {
        u32 dim;
        f32 birth, death;
    };
    std::vector<Bar> bars;
    for (u32 d = 0; d < bc.birth.size(); ++d)
        for (u64 i = 0; i < bc.birth[d].size(); ++i) bars.push_back({d, bc.birth[d][i], bc.death[d][i]});
    std::stable_sort(bars.begin(), bars.end(), [](const Bar& a, const Bar& b) {
        if (a.dim != b.dim) return a.dim < b.dim;
        if (a.birth != b.birth) return a.birth < b.birth;
        return a.death < b.death;
    });
    for (const Bar& b : bars) {
        if (std::isinf(b.death)) std::fprintf(f, "%u %.9g inf\n", b.dim, double(b.birth));
        else std::fprintf(f, "%u %.9g %.9g\n", b.dim, double(b.birth), double(b.death));
    }
}

void print_summary(const Result& r) {
    const Stats& s = r.stats;
    std::fprintf(stderr, "turbotops: mode: %s\n", s.mode_reason.c_str());
    std::fprintf(stderr,
                 "turbotops: method=%s threshold=%g edges=%llu index_bits=%d graph=%.3fs total=%.3fs "
                 "peak_rss=%.1fMB\n",
                 s.method.c_str(), double(s.threshold), (unsigned long long)s.edges, s.wide_index ? 128 : 64,
                 s.seconds_graph, s.seconds_total, s.peak_rss_mb);
    if (s.engine == "coboundary" && !s.columns.empty()) {
        // per dimension d >= 1: columns reduced, and columns skipped as
        // apparent during assembly or closed by the emergent shortcut
        std::fprintf(stderr, "turbotops: coboundary telemetry:");
        for (size_t d = 0; d < s.columns.size(); ++d) {
            const u64 skipped = (d + 1 < s.apparent.size() ? s.apparent[d + 1] : 0) +
                                (d < s.emergent.size() ? s.emergent[d] : 0);
            std::fprintf(stderr, " H%zu columns=%llu skipped=%llu", d + 1, (unsigned long long)s.columns[d],
                         (unsigned long long)skipped);
        }
        std::fputc('\n', stderr);
    }
    for (u32 d = 0; d < r.barcode.birth.size(); ++d) {
        u64 finite = 0, essential = 0;
        for (f32 e : r.barcode.death[d]) (std::isinf(e) ? essential : finite) += 1;
        std::fprintf(stderr, "turbotops: H%u %llu finite, %llu essential\n", d, (unsigned long long)finite,
                     (unsigned long long)essential);
    }
}

struct RunArgs {
    Options opt;
    std::string input, output;
    u32 dim = 0;
    bool edges = false, quiet = false, verify = false;
    bool mode_set = false;
};

void set_mode(RunArgs& a, Mode m, const std::string& flag) {
    if (a.mode_set && a.opt.mode != m) fail(flag + " conflicts with an earlier mode/engine flag");
    a.mode_set = true;
    a.opt.mode = m;
}

Mode engine_mode(const std::string& e) {
    if (e == "auto") return Mode::Auto;
    if (e == "reducer") return Mode::Coboundary;
    if (e == "redzed") return Mode::Boundary;
    throw std::invalid_argument("engine must be auto, reducer or redzed");
}

RunArgs parse_run(const std::vector<std::string>& args) {
    RunArgs a;
    u32 unused_hnsw = 0;
    for (size_t i = 1; i < args.size(); ++i) {
        const std::string& f = args[i];
        auto value = [&]() -> const std::string& {
            if (i + 1 >= args.size()) fail("missing value for " + f);
            return args[++i];
        };
        if (f == "--maxdim") a.opt.maxdim = to_u32(value(), "maxdim");
        else if (f == "--threshold") a.opt.threshold = to_f32(value(), "threshold");
        else if (f == "--metric") a.opt.metric = value();
        else if (f == "--mode") set_mode(a, parse_mode(value()), f);
        else if (f == "--coboundary") set_mode(a, Mode::Coboundary, f);
        else if (f == "--boundary" || f == "--no-coboundary") set_mode(a, Mode::Boundary, f);
        else if (f == "--engine") set_mode(a, engine_mode(value()), f);
        else if (f == "--graph") {
            const std::string& g = value();
            if (g == "full") a.opt.graph = Options::Graph::Full;
            else if (g == "knn") a.opt.graph = Options::Graph::Knn;
            else fail("--graph must be full or knn");
        } else if (f == "-k") a.opt.k = to_u32(value(), "k");
        else if (f == "--approximate") a.opt.approximate = true;
        else if (f == "--hnsw-m") unused_hnsw = to_u32(value(), "M");
        else if (f == "--ef-construction" || f == "--ef-search") unused_hnsw = to_u32(value(), "ef");
        else if (f == "--verify-recall") {
            a.verify = true;
            a.opt.recall_samples = to_u64(value(), "samples");
        } else if (f == "--dim") a.dim = to_u32(value(), "dim");
        else if (f == "--threads") a.opt.threads = to_u32(value(), "threads");
        else if (f == "--memory-budget") a.opt.memory_budget = to_bytes(value());
        else if (f == "--output") a.output = value();
        else if (f == "--edges") a.edges = true;
        else if (f == "--quiet") a.quiet = true;
        else if (!f.empty() && f[0] == '-') fail("unknown option '" + f + "'");
        else if (a.input.empty()) a.input = f;
        else fail("unexpected argument '" + f + "'");
    }
    (void)unused_hnsw;
    if (a.input.empty()) fail("run: missing input file");
    if (a.verify && !a.opt.approximate) fail("--verify-recall requires --approximate");
    if (a.verify && a.opt.graph != Options::Graph::Knn) fail("--verify-recall requires --graph knn");
    if (a.opt.approximate && a.opt.graph != Options::Graph::Knn) fail("--approximate requires --graph knn");
    return a;
}

int run(const std::vector<std::string>& args) {
    RunArgs a = parse_run(args);
    // open the output first: a bad path must not cost the computation
    FILE* out = stdout;
    if (!a.output.empty()) {
        out = std::fopen(a.output.c_str(), "w");
        if (out == nullptr) fail("cannot write " + a.output);
    }
    Result r;
    if (a.edges) {
        const CsrGraph g = load_edges(a.input);
        if (g.n == 0) fail(a.input + ": no vertices (empty edge list?)");
        r = persistence(g, a.opt.maxdim, a.opt.threshold, a.opt.threads, a.opt.mode);
    } else {
        u32 n = 0, dim = a.dim;
        const std::vector<f32> x = has_suffix(a.input, ".bin") ? load_bin(a.input, dim, n) : load_text(a.input, n, dim);
        if (n == 0) fail(a.input + ": no points (empty file, directory, or wrong format?)");
        if (a.verify && a.opt.recall_samples == 0) a.opt.recall_samples = n;
        r = persistence(x.data(), n, dim, a.opt);
    }
    print_diagram(out, r.barcode);
    if (out != stdout) std::fclose(out);
    if (!a.quiet) print_summary(r);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::vector<std::string> args(argv + 1, argv + argc);
    if (args.empty() || args[0] == "-h" || args[0] == "--help" || args[0] == "help") {
        std::fputs(kHelp, args.empty() ? stderr : stdout);
        return args.empty() ? 2 : 0;
    }
    if (args[0] == "--version") {
        std::printf("turbotops 0.1.0 (MI355X front end)\n");
        return 0;
    }
    try {
        if (args[0] == "gen" || args[0] == "test") fail("'" + args[0] + "' is not part of this tool (only 'run')");
        if (args[0] != "run") fail("unknown command '" + args[0] + "' (try --help)");
        return run(args);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "turbotops: %s\n", e.what());
        return 2;
    }
}
