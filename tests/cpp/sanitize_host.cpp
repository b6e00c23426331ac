// The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer:
// the proxy ABI of gsea_proxy.cpp (symbol -> CSR mapping, the C++ proxies,
// dump naming and writing, the C entry gsea_map_sets) and the GCT/GMT/CLS
// loaders of gsea_io.cpp, on the reference's own GMT and CLS files and on
// files written here.  The engine entry points that host code calls are
// answered on the CPU by the oracle (oracle/gsea_oracle.c, also built with
// the sanitizers), so no GPU is involved; the loaders and proxies are the
// library's own sources, compiled unchanged.
// Usage: sanitize_host <golden dir> <scratch dir>.  Exit status 0 when every
// check holds; prints one line per failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cudaGSEA_wrapper.cuh"
#include "gsea_hip.h"
#include "gsea_oracle.h"

static int failures = 0;
static std::string g_err;

static void expect(bool ok, const std::string &what) {
    if (!ok) {
        ++failures;
        std::cerr << "FAIL: " << what << std::endl;
    }
}

// ---- the engine entry points gsea_proxy.cpp calls, answered by the oracle
namespace gsea {
int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
}  // namespace gsea

extern "C" {
const char *gsea_last_error(void) { return g_err.c_str(); }
int gsea_metric_id(const char *name) { return orc_metric_id(name); }
int gsea_run_f32(const float *x, int64_t n, int64_t s, const uint8_t *labels,
                 const int32_t *off, const int32_t *genes, int64_t g, int64_t p, int metric,
                 int flags, float *es_out, float *stats_out, gsea_timings *) {
    return orc_gsea_f32(x, n, s, labels, off, genes, g, p, metric, flags, 1, es_out,
                        stats_out);
}
int gsea_run_f64(const double *x, int64_t n, int64_t s, const uint8_t *labels,
                 const int32_t *off, const int32_t *genes, int64_t g, int64_t p, int metric,
                 int flags, double *es_out, double *stats_out, gsea_timings *) {
    return orc_gsea_f64(x, n, s, labels, off, genes, g, p, metric, flags, 1, es_out,
                        stats_out);
}
int gsea_device_count(int *count) {
    *count = 1;
    return 0;
}
int gsea_get_device(int *device) {
    *device = 0;
    return 0;
}
int gsea_set_device(int device) { return device == 0 ? 0 : GSEA_ERR_INVALID_ARGUMENT; }
int gsea_device_name(int, char *name, size_t len, size_t *total) {
    std::snprintf(name, len, "%s", "cpu answer");
    *total = (size_t)3 << 30;
    return 0;
}
}

static bool throws(void (*f)(const std::string &), const std::string &arg) {
    try {
        f(arg);
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

static long file_size(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    return f ? (long)f.tellg() : -1;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: sanitize_host <golden dir> <scratch dir>" << std::endl;
        return 2;
    }
    const std::string gold = argv[1], tmp = argv[2];

    // the reference's GMT and CLS files
    std::vector<std::string> pname;
    std::vector<std::vector<std::string>> pathw;
    loadGeneSetsFromGMT_proxy(gold + "/h.all.v5.0.symbols.gmt", pname, pathw);
    expect(pname.size() == 50 && pathw.size() == 50, "hallmark GMT: 50 sets");
    std::vector<proxy_label_t> cls;
    loadLabelsFromCLS_proxy(gold + "/GSE19429_series.cls", cls);
    bool binary = !cls.empty();
    for (auto l : cls) binary = binary && l <= 1;
    expect(binary, "GSE19429 CLS: labels 0/1");

    // a GCT over the first 400 symbols of the hallmark sets plus 40 others,
    // 12 samples
    std::vector<std::string> sym;
    for (auto &set : pathw)
        for (auto &x : set)
            if (sym.size() < 400 && std::find(sym.begin(), sym.end(), x) == sym.end())
                sym.push_back(x);
    for (int i = 0; i < 40; ++i) sym.push_back("EXTRA" + std::to_string(i));
    const int S = 12;
    {
        std::ofstream f(tmp + "/x.gct");
        f << "#1.2\n" << sym.size() << "\t" << S << "\nNAME\tDescription";
        for (int s = 0; s < S; ++s) f << "\tS" << s;
        f << "\n";
        for (size_t g = 0; g < sym.size(); ++g) {
            f << " " << sym[g] << " \tna";
            for (int s = 0; s < S; ++s) f << "\t" << 5.0 + std::sin(1.0 + g * 0.37 + s * 1.3);
            f << "\n";
        }
    }
    std::vector<proxy_value_t> ex;
    std::vector<std::string> gsymb, plist;
    loadExpressionDataFromGCT_proxy(tmp + "/x.gct", ex, gsymb, plist);
    expect(gsymb.size() == sym.size() && plist.size() == (size_t)S &&
               ex.size() == sym.size() * S,
           "GCT: sizes");
    expect(!gsymb.empty() && gsymb[0] == sym[0], "GCT: symbols trimmed");

    // broken inputs raise, never end the process
    expect(throws([](const std::string &p) {
               std::vector<proxy_value_t> e;
               std::vector<std::string> a, b;
               loadExpressionDataFromGCT_proxy(p, e, a, b);
           }, tmp + "/absent.gct"),
           "missing GCT raises");
    {
        std::ofstream f(tmp + "/bad.gct");
        f << "#1.3\n2\t2\nNAME\tDescription\tA\tB\ng1\tna\t1\t2\n";
    }
    expect(throws([](const std::string &p) {
               std::vector<proxy_value_t> e;
               std::vector<std::string> a, b;
               loadExpressionDataFromGCT_proxy(p, e, a, b);
           }, tmp + "/bad.gct"),
           "bad GCT header raises");
    {
        std::ofstream f(tmp + "/bad.cls");
        f << "4 2 1\n# A B\n0 1 1\n";
    }
    expect(throws([](const std::string &p) {
               std::vector<proxy_label_t> l;
               loadLabelsFromCLS_proxy(p, l);
           }, tmp + "/bad.cls"),
           "short CLS raises");

    // the C++ proxies, fp32 with a dump and fp64 with swapped labels
    std::vector<char> lab(S);
    for (int s = 0; s < S; ++s) lab[s] = (char)(s % 3 == 0 ? 1 : 0);
    std::vector<float> ex32(ex.begin(), ex.end());
    const size_t P = 64;
    const std::string dump = tmp + "/run";
    std::vector<float> r32 = gsea_single_proxy(ex32, gsymb, plist, lab, pname, pathw,
                                               "onepass_t_test", P, true, false, dump);
    expect(r32.size() == 5 * pname.size(), "fp32 proxy: 5G results");
    bool bounded = true;
    for (size_t i = 0; i < pname.size(); ++i)
        bounded = bounded && (std::isnan(r32[i]) || std::fabs(r32[i]) <= 1.0f);
    expect(bounded, "fp32 proxy: |ES| <= 1");
    const std::string name = dump + "_" + std::to_string(pname.size()) + "_" +
                             std::to_string(P) + "_32.es";
    expect(file_size(name) == (long)(pname.size() * P * 4), "fp32 dump: name and size");
    std::vector<char> lab2 = lab;
    std::vector<double> r64 = gsea_double_proxy(ex, gsymb, plist, lab2, pname, pathw,
                                                "stable_signal2noise", P, false, true, "");
    expect(r64.size() == 5 * pname.size(), "fp64 proxy: 5G results");
    bool flipped = true;
    for (int s = 0; s < S; ++s) flipped = flipped && lab2[s] == (char)(1 - lab[s]);
    expect(flipped, "fp64 proxy: labels swapped in place");
    bool raised = false;
    try {
        gsea_single_proxy(ex32, gsymb, plist, lab, pname, pathw, "no_such_metric", P, true,
                          false, "");
    } catch (const std::exception &) {
        raised = true;
    }
    expect(raised, "unknown metric raises");

    // the C entry: duplicates collapse, absent symbols are ignored
    const char *genes[] = {"A", "B", "C", "B"};
    const char *setsym[] = {"B", "B", "Z", "A", "C", "Q"};
    const int64_t setoff[] = {0, 3, 6};
    int32_t off[3];
    int64_t total = -1;
    // gene symbol B occurs twice in the data (genes 1 and 3): set 0 = {B, Z}
    // -> genes 1, 3; set 1 = {A, C, Q} -> genes 0, 2
    int rc = gsea_map_sets(genes, 4, setsym, setoff, 2, off, nullptr, 0, &total);
    expect(rc == 0 && total == 4, "map_sets: size query");
    std::vector<int32_t> members(4, -1);
    rc = gsea_map_sets(genes, 4, setsym, setoff, 2, off, members.data(), 4, &total);
    expect(rc == 0 && off[0] == 0 && off[1] == 2 && off[2] == 4, "map_sets: offsets");
    expect(members == std::vector<int32_t>({1, 3, 0, 2}), "map_sets: members");

    if (failures == 0) std::cout << "sanitized host checks passed" << std::endl;
    return failures == 0 ? 0 : 1;
}
