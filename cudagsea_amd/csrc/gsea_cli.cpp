// gsea_cli.cpp -- the standalone command line of cudaGSEA/src/cudaGSEA.cu:8-191
// (flags of cudaGSEA/src/HELP.txt) on top of libgsea_hip.so: reads a GCT
// expression table, a CLS label file and a GMT gene-set file with the loader
// proxies, runs permutation GSEA on the GPU through the C++ proxy ABI and
// prints one RESULT line per gene set in the reference's format
// (cudaGSEA/src/include/cugsea.cuh:248-255).
//
// Differences: errors print "ERROR: ..." and return the reference's error code
// instead of calling exit() from deep inside the run; -cpu is refused (this
// build has no CPU path); -parse-only reads and checks the three files and
// prints their sizes without touching the GPU.
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cudaGSEA_wrapper.cuh"
#include "gsea_error.hpp"
#include "gsea_hip.h"

namespace {

// exit statuses of cudaGSEA/src/cudaGSEA.cu (error_codes.cuh numbers; a
// missing -cls exits with the no-GMT code there, cudaGSEA.cu:52-57)
enum {
    kIncompatibleGctCls = 9, kNoGct = 11, kNoGmt = 13, kCannotResolveGct = 14,
    kCannotResolveCls = 15, kCannotResolveGmt = 16, kCannotParseInteger = 17,
    kInvalidOrder = 18, kEmptyDump = 19, kInvalidPrecision = 20, kGpuAndCpu = 22,
};

void print_help() {
    std::cout <<
        "INFO: Please, use cudagsea as follows:\n"
        "INFO: Note, flags in square brackets [...] are optional. Entities in angle\n"
        "INFO: brackets <...> denote the default settings when flag not specified.\n"
        "INFO:\n"
        "INFO: cudagsea -res /path/to/collapsed/gene/expression/file.gct            \\\n"
        "INFO:          -cls /path/to/class/label/assignment/file.cls               \\\n"
        "INFO:          -gmx /path/to/gene/set/collection/file                      \\\n"
        "INFO:         [-nperm number_of_permutations (positive integer <16384>)]   \\\n"
        "INFO:         [-metric metric_string (see README.md <onepass_signal2noise>)]\\\n"
        "INFO:         [-order order_predicate (ascending/descending <descending>)] \\\n"
        "INFO:         [-dump /path/to/binary/dump/file <no file dumped if left out>]\\\n"
        "INFO:         [-precision floating_point_precision (single/double <single>)]\\\n"
        "INFO:         [-gpu gpu_id (positive integer) <0>]                         \\\n"
        "INFO:         [-parse-only]\n";
}

bool has(int argc, char **argv, const std::string &flag) {
    for (int i = 1; i < argc; ++i)
        if (flag == argv[i]) return true;
    return false;
}

std::string value(int argc, char **argv, const std::string &flag) {
    for (int i = 1; i + 1 < argc; ++i)
        if (flag == argv[i]) return argv[i + 1];
    return "";
}

bool parse_count(const std::string &s, long long *out) {
    if (s.empty()) return false;
    for (char c : s)
        if (c < '0' || c > '9') return false;
    errno = 0;
    const long long v = std::strtoll(s.c_str(), nullptr, 10);
    if (errno) return false;
    *out = v;
    return true;
}

std::string resolve(const std::string &p) {
    if (p.empty()) return "";
    char buf[PATH_MAX];
    return realpath(p.c_str(), buf) ? std::string(buf) : std::string();
}

int fail(int code, const std::string &msg) {
    std::cout << "ERROR: " << msg << " Exiting." << std::endl;
    return code;
}

std::vector<float> proxy(std::vector<float> &x, std::vector<std::string> &genes,
                         std::vector<std::string> &samples, std::vector<char> &labels,
                         std::vector<std::string> &names,
                         std::vector<std::vector<std::string>> &sets,
                         const std::string &metric, size_t nperm, bool descending,
                         const std::string &dump) {
    return gsea_single_proxy(x, genes, samples, labels, names, sets, metric, nperm,
                             descending, false, dump);
}

std::vector<double> proxy(std::vector<double> &x, std::vector<std::string> &genes,
                          std::vector<std::string> &samples, std::vector<char> &labels,
                          std::vector<std::string> &names,
                          std::vector<std::vector<std::string>> &sets,
                          const std::string &metric, size_t nperm, bool descending,
                          const std::string &dump) {
    return gsea_double_proxy(x, genes, samples, labels, names, sets, metric, nperm,
                             descending, false, dump);
}

template <class T>
int run(const std::vector<double> &gene_major, size_t N, size_t S,
        std::vector<std::string> &genes, std::vector<std::string> &samples,
        std::vector<char> &labels, std::vector<std::string> &names,
        std::vector<std::vector<std::string>> &sets, const std::string &metric,
        size_t nperm, bool descending, const std::string &dump) {
    // GCT rows are genes; the engine takes sample-major [S][N]
    std::vector<T> exprs(N * S);
    for (size_t g = 0; g < N; ++g)
        for (size_t s = 0; s < S; ++s) exprs[s * N + g] = (T)gene_major[g * S + s];
    const std::vector<T> res = proxy(exprs, genes, samples, labels, names, sets, metric,
                                     nperm, descending, dump);
    const size_t G = names.size();
    for (size_t p = 0; p < G; ++p) {
        std::cout << std::showpos << std::fixed << "RESULT:"
                  << " ES: " << res[p] << " NES: " << res[G + p]
                  << " NP: " << res[2 * G + p] << " FWER: " << res[3 * G + p]
                  << "\t(" << names[p] << ")" << std::endl;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    std::cout << "INFO: " << gsea_version() << " -- permutation GSEA on AMD Instinct"
              << std::endl;
    if (argc == 1) {
        print_help();
        return 1;
    }
    if (!has(argc, argv, "-res")) {
        print_help();
        return fail(kNoGct, "You have to specify a collapsed gene expression file with -res.");
    }
    if (!has(argc, argv, "-cls")) {
        print_help();
        return fail(kNoGmt, "You have to specify a class label assignment file with -cls.");
    }
    const std::string gct = resolve(value(argc, argv, "-res"));
    if (gct.empty())
        return fail(kCannotResolveGct,
                    "File specified with -res could not be resolved by the operating system.");
    const std::string cls = resolve(value(argc, argv, "-cls"));
    if (cls.empty())
        return fail(kCannotResolveCls,
                    "File specified with -cls could not be resolved by the operating system.");
    const std::string gmt = resolve(value(argc, argv, "-gmx"));
    if (gmt.empty())
        return fail(kCannotResolveGmt,
                    "File specified with -gmx could not be resolved by the operating system.");

    long long nperm = 1 << 14;
    if (has(argc, argv, "-nperm") && !parse_count(value(argc, argv, "-nperm"), &nperm))
        return fail(kCannotParseInteger,
                    "number specified with -nperm could not be parsed as integer.");
    std::string metric = "onepass_signal2noise";
    if (has(argc, argv, "-metric")) metric = value(argc, argv, "-metric");
    bool descending = true;
    if (has(argc, argv, "-order")) {
        const std::string o = value(argc, argv, "-order");
        if (o == "ascending") descending = false;
        else if (o != "descending")
            return fail(kInvalidOrder,
                        "option specified with -order must be either ascending or descending.");
    }
    std::string dump;
    if (has(argc, argv, "-dump")) {
        dump = value(argc, argv, "-dump");
        if (dump.empty())
            return fail(kEmptyDump, "Parameter -dump observed but empty file name found.");
    }
    bool single = true;
    if (has(argc, argv, "-precision")) {
        const std::string p = value(argc, argv, "-precision");
        if (p == "double") single = false;
        else if (p != "single")
            return fail(kInvalidPrecision,
                        "option specified with -precision must be either single or double.");
    }
    if (has(argc, argv, "-gpu") && has(argc, argv, "-cpu"))
        return fail(kGpuAndCpu, "you can only specify -cpu if -gpu is not already set.");
    if (has(argc, argv, "-cpu"))
        return fail(kGpuAndCpu, "this build runs on the GPU only; -cpu is not provided.");
    if (has(argc, argv, "-gpu")) {
        long long dev = 0;
        if (!parse_count(value(argc, argv, "-gpu"), &dev))
            return fail(kCannotParseInteger,
                        "number specified with -gpu could not be parsed as integer.");
        if (!has(argc, argv, "-parse-only")) setCudaDevice_proxy((int)dev);
    }

    try {
        std::vector<double> gene_major;
        std::vector<std::string> genes, samples, names;
        std::vector<size_t> lab;
        std::vector<std::vector<std::string>> sets;
        loadExpressionDataFromGCT_proxy(gct, gene_major, genes, samples);
        loadLabelsFromCLS_proxy(cls, lab);
        loadGeneSetsFromGMT_proxy(gmt, names, sets);
        const size_t N = genes.size(), S = samples.size();
        if (lab.size() != S)
            return fail(kIncompatibleGctCls,
                        "Incompatible class file (cls) and expression data file.");
        size_t nA = 0;
        for (size_t v : lab) nA += v == 0;
        std::cout << "INFO: expression data: " << gct << " (" << N << " genes x " << S
                  << " samples)" << std::endl;
        std::cout << "INFO: phenotype labels: " << cls << " (" << nA << " / " << S - nA
                  << ")" << std::endl;
        std::cout << "INFO: gene sets: " << gmt << " (" << names.size() << " sets)"
                  << std::endl;
        if (has(argc, argv, "-parse-only")) return 0;
        if (nperm <= 0)
            return fail(kCannotParseInteger, "-nperm must be a positive integer.");
        std::vector<char> labels(lab.begin(), lab.end());
        return single ? run<float>(gene_major, N, S, genes, samples, labels, names, sets,
                                   metric, (size_t)nperm, descending, dump)
                      : run<double>(gene_major, N, S, genes, samples, labels, names, sets,
                                    metric, (size_t)nperm, descending, dump);
    } catch (const gsea::Error &e) {
        std::cout << "ERROR: " << e.what() << std::endl;
        return e.code;
    } catch (const std::exception &e) {
        std::cout << "ERROR: " << e.what() << std::endl;
        return 1;
    }
}
