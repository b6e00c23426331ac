// gsea_io.cpp -- Broad file formats (GCT #1.2, GMT, CLS) for the loader
// proxies of cudaGSEA/src/cudaGSEA_wrapper.cuh:15-25.  Semantics follow
// cudaGSEA/src/include/read_data_files.cuh:46-321 (header checks, symbol
// trimming, duplicate-symbol rejection, symbolic or 0/1 class labels); where
// the reference calls exit() with an error code these functions throw
// gsea::Error (a std::runtime_error) carrying the same code, which Rcpp turns
// into an R error instead of ending the R session.
#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "cudaGSEA_wrapper.cuh"
#include "gsea_error.hpp"
#include "gsea_hip.h"

namespace {

using IoError = gsea::Error;

std::string trim(std::string s) {
    auto sp = [](unsigned char c) { return std::isspace(c) != 0; };
    s.erase(s.begin(), std::find_if(s.begin(), s.end(), [&](char c) { return !sp(c); }));
    s.erase(std::find_if(s.rbegin(), s.rend(), [&](char c) { return !sp(c); }).base(), s.end());
    return s;
}

// read_data_files.cuh:46-127; exprs is gene-major [N][S] as in the GCT file
void read_gct(const std::string &path, std::vector<double> &exprs,
              std::vector<std::string> &samples, std::vector<std::string> &genes) {
    std::ifstream f(path);
    if (!f) throw IoError(2, "Unable to open file " + path);       // CANNOT_OPEN_FILE
    std::string tok;
    std::getline(f, tok);
    if (trim(tok) != "#1.2") throw IoError(4, path + " is not a valid gct header");
    size_t N = 0, S = 0;
    f >> N >> S;
    if (!f || S == 0) throw IoError(4, path + " has no valid dimension line");
    std::getline(f, tok);                                          // end of line
    std::string header;
    std::getline(f, header);
    std::stringstream hs(header);
    std::getline(hs, tok, '\t');                                   // NAME
    std::getline(hs, tok, '\t');                                   // Description
    samples.assign(S, "");
    for (size_t s = 0; s < S; ++s) {
        std::getline(hs, tok, '\t');
        samples[s] = trim(tok);
    }
    exprs.assign(N * S, 0.0);
    genes.assign(N, "");
    std::unordered_set<std::string> seen;
    for (size_t g = 0; g < N; ++g) {
        std::getline(f, tok, '\t');
        genes[g] = trim(tok);
        if (!seen.insert(genes[g]).second)
            throw IoError(3, "duplicate symbol " + genes[g]);      // DUPLICATE_GENE_SYMBOL
        std::getline(f, tok, '\t');                                // description
        for (size_t s = 0; s < S; ++s) f >> exprs[g * S + s];
        if (!f) throw IoError(4, path + ": truncated expression table");
    }
}

// read_data_files.cuh:133-187
void read_gmt(const std::string &path, std::vector<std::string> &names,
              std::vector<std::vector<std::string>> &sets) {
    std::ifstream f(path);
    if (!f) throw IoError(2, "Unable to open file " + path);
    names.clear();
    sets.clear();
    std::string line, tok;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::stringstream ss(line);
        std::vector<std::string> syms;
        std::getline(ss, tok, '\t');
        names.push_back(tok);
        size_t count = 0;
        while (std::getline(ss, tok, '\t'))
            if (count++) syms.push_back(trim(tok));                // skip description
        sets.push_back(std::move(syms));
    }
}

// read_data_files.cuh:193-321
void read_cls(const std::string &path, std::vector<size_t> &labels) {
    std::ifstream f(path);
    if (!f) throw IoError(2, "Unable to open file " + path);
    std::string line;
    std::getline(f, line);
    std::stringstream ss(line);
    size_t S = 0, classes = 0, one = 0;
    ss >> S >> classes >> one;
    if (classes != 2) throw IoError(6, "cls file must name exactly two classes");
    if (one != 1) throw IoError(5, "invalid cls header in " + path);
    std::getline(f, line);
    std::stringstream ss2(line);
    std::string hash, c1, c2;
    std::getline(ss2, hash, '#');
    ss2 >> c1 >> c2;
    c1 = trim(c1);
    c2 = trim(c2);
    std::getline(f, line);
    std::stringstream ss3(line);
    labels.assign(S, 0);
    size_t a = 0, b = 0;
    for (size_t i = 0; i < S; ++i) {
        std::string x;
        ss3 >> x;
        if (x == c1 || x == "0") labels[i] = 0;
        else if (x == c2 || x == "1") labels[i] = 1;
        else throw IoError(7, "Unknown label in cls file " + path + " : \"" + x + "\"");
        (labels[i] ? b : a) += 1;
    }
    if (a + b != S) throw IoError(8, "corrupted cls file " + path);
}

}  // namespace

// ---- C++ proxy ABI (cudaGSEA_wrapper.cuh:15-25) ----

void loadLabelsFromCLS_proxy(std::string path, std::vector<proxy_label_t> &labels) {
    read_cls(path, labels);
}

void loadGeneSetsFromGMT_proxy(std::string path, std::vector<proxy_strng_t> &names,
                               std::vector<std::vector<proxy_strng_t>> &sets) {
    read_gmt(path, names, sets);
}

void loadExpressionDataFromGCT_proxy(std::string path, std::vector<proxy_value_t> &exprs,
                                     std::vector<proxy_strng_t> &genes,
                                     std::vector<proxy_strng_t> &samples) {
    read_gct(path, exprs, samples, genes);
}
