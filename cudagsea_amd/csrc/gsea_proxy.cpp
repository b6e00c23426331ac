// gsea_proxy.cpp -- symbol-level host layer above the C ABI:
//   * gsea_map_sets: gene-set symbols -> CSR of gene indices
//     (create_bitmaps semantics, cudaGSEA/src/include/create_bitmaps.cuh:13-65)
//   * gsea_dump_es: the reference's raw .es dump (dump_binary.cuh:10-30,
//     naming cudaGSEA_wrapper.cu:384-390)
//   * gsea_proxy_run: one R-level GSEA call with symbols
//   * the C++ proxy ABI of cudaGSEA/src/cudaGSEA_wrapper.cuh:15-61
//     (gsea_single_proxy, gsea_double_proxy, device proxies), so the Rcpp
//     glue cudaGSEA_wrapper.cpp links against this library unchanged.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "cudaGSEA_wrapper.cuh"
#include "gsea_hip.h"

namespace gsea {
int set_error(int code, const std::string &msg);   // gsea_engine.hip
}

namespace {

int proxy_fail(int code, const std::string &msg) { return gsea::set_error(code, msg); }

// CSR build shared by the C entry and the C++ proxies
struct SetCSR {
    std::vector<int32_t> off;
    std::vector<int32_t> genes;
};

template <class GetGene, class GetSetSym>
SetCSR map_sets_impl(int64_t N, GetGene gene_sym, int64_t G,
                     const int64_t *sym_off, GetSetSym set_sym) {
    std::unordered_map<std::string, std::vector<int32_t>> where;
    where.reserve((size_t)N * 2);
    for (int64_t g = 0; g < N; ++g) where[gene_sym(g)].push_back((int32_t)g);
    SetCSR csr;
    csr.off.assign((size_t)G + 1, 0);
    std::vector<int32_t> members;
    std::unordered_set<std::string> seen;
    for (int64_t s = 0; s < G; ++s) {
        members.clear();
        seen.clear();
        for (int64_t i = sym_off[s]; i < sym_off[s + 1]; ++i) {
            const std::string sym = set_sym(i);
            if (!seen.insert(sym).second) continue;
            auto it = where.find(sym);
            if (it == where.end()) continue;
            members.insert(members.end(), it->second.begin(), it->second.end());
        }
        std::sort(members.begin(), members.end());
        members.erase(std::unique(members.begin(), members.end()), members.end());
        csr.genes.insert(csr.genes.end(), members.begin(), members.end());
        csr.off[(size_t)s + 1] = (int32_t)csr.genes.size();
    }
    return csr;
}

std::string dump_name(const std::string &prefix, int64_t G, int64_t P, int bits) {
    return prefix + "_" + std::to_string(G) + "_" + std::to_string(P) + "_" +
           std::to_string(bits) + ".es";
}

template <class T>
int proxy_run_t(const T *exprs, int64_t N, int64_t S, const uint8_t *labels,
                const SetCSR &csr, int64_t G, int metric, int64_t P, int flags,
                const std::string &dump, T *stats, gsea_timings *t) {
    std::vector<T> es;
    if (!dump.empty()) es.resize((size_t)G * P);
    int rc;
    if (sizeof(T) == 4)
        rc = gsea_run_f32((const float *)exprs, N, S, labels, csr.off.data(),
                          csr.genes.data(), G, P, metric, flags,
                          dump.empty() ? nullptr : (float *)es.data(),
                          (float *)stats, t);
    else
        rc = gsea_run_f64((const double *)exprs, N, S, labels, csr.off.data(),
                          csr.genes.data(), G, P, metric, flags,
                          dump.empty() ? nullptr : (double *)es.data(),
                          (double *)stats, t);
    if (rc) return rc;
    if (!dump.empty()) {
        rc = gsea_dump_es(dump.c_str(), es.data(), G, P, (int)(8 * sizeof(T)));
        if (rc) return rc;
    }
    return GSEA_OK;
}

template <class T>
std::vector<T> cxx_proxy(std::vector<T> &exprs, std::vector<std::string> &gsymb,
                         std::vector<std::string> &plist, std::vector<char> &labels,
                         std::vector<std::string> &pname,
                         std::vector<std::vector<std::string>> &pathw,
                         std::string metric, size_t num_perms, bool sort_direction,
                         bool swap_labels, std::string dump_filename) {
    const int64_t N = (int64_t)gsymb.size();
    const int64_t S = (int64_t)plist.size();
    const int64_t G = (int64_t)pname.size();
    const int mid = gsea_metric_id(metric.c_str());
    if (mid < 0) throw std::runtime_error("no known metric specified: " + metric);
    if ((int64_t)exprs.size() != N * S || (int64_t)labels.size() != S ||
        (int64_t)pathw.size() < G)
        throw std::runtime_error("inconsistent input sizes");
    std::vector<int64_t> sym_off((size_t)G + 1, 0);
    for (int64_t s = 0; s < G; ++s) sym_off[(size_t)s + 1] = sym_off[(size_t)s] + (int64_t)pathw[(size_t)s].size();
    std::vector<const std::string *> flat;
    flat.reserve((size_t)sym_off[(size_t)G]);
    for (int64_t s = 0; s < G; ++s)
        for (auto &x : pathw[(size_t)s]) flat.push_back(&x);
    SetCSR csr = map_sets_impl(
        N, [&](int64_t g) { return gsymb[(size_t)g]; }, G, sym_off.data(),
        [&](int64_t i) { return *flat[(size_t)i]; });
    std::vector<uint8_t> lab(labels.begin(), labels.end());
    // the reference proxy flips the caller's labels in place
    // (cudaGSEA_wrapper.cu:246-251); keep that visible side effect
    if (swap_labels)
        for (auto &l : labels) l = (char)((l + 1) & 1);
    const int flags = (sort_direction ? 0 : GSEA_FLAG_ASCENDING) |
                      (swap_labels ? GSEA_FLAG_SWAP_LABELS : 0);
    std::vector<T> stats((size_t)5 * G);
    const int rc = proxy_run_t<T>(exprs.data(), N, S, lab.data(), csr, G, mid,
                                  (int64_t)num_perms, flags, dump_filename,
                                  stats.data(), nullptr);
    if (rc) throw std::runtime_error(std::string("cudaGSEA engine error: ") + gsea_last_error());
    return stats;
}

}  // namespace

extern "C" {

int gsea_map_sets(const char *const *gene_symbols, int64_t N,
                  const char *const *set_symbols, const int64_t *set_symbol_offsets,
                  int64_t G, int32_t *set_offsets_out, int32_t *set_genes_out,
                  int64_t capacity, int64_t *total_out) {
    if (N < 0 || G < 0 || (N && !gene_symbols) || !set_symbol_offsets ||
        !set_offsets_out || !total_out)
        return proxy_fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    SetCSR csr = map_sets_impl(
        N, [&](int64_t g) { return std::string(gene_symbols[g]); }, G,
        set_symbol_offsets, [&](int64_t i) { return std::string(set_symbols[i]); });
    *total_out = (int64_t)csr.genes.size();
    std::memcpy(set_offsets_out, csr.off.data(), sizeof(int32_t) * (size_t)(G + 1));
    if (capacity >= (int64_t)csr.genes.size() && set_genes_out)
        std::memcpy(set_genes_out, csr.genes.data(), sizeof(int32_t) * csr.genes.size());
    return GSEA_OK;
}

int gsea_dump_es(const char *prefix, const void *es, int64_t G, int64_t P, int bits) {
    if (!prefix || !*prefix)
        return proxy_fail(GSEA_ERR_EMPTY_DUMP_FILE_NAME, "empty dump file name");
    if (!es || (bits != 32 && bits != 64))
        return proxy_fail(GSEA_ERR_INVALID_ARGUMENT, "bad dump arguments");
    const std::string name = dump_name(prefix, G, P, bits);
    std::ofstream f(name.c_str(), std::ios::binary);
    if (!f) return proxy_fail(GSEA_ERR_CANNOT_OPEN_FILE, "cannot open " + name);
    f.write((const char *)es, (std::streamsize)((size_t)G * P * (bits / 8)));
    if (!f) return proxy_fail(GSEA_ERR_CANNOT_OPEN_FILE, "cannot write " + name);
    return GSEA_OK;
}

int gsea_proxy_run(int double_precision, const void *exprs, int64_t N, int64_t S,
                   const char *const *gene_symbols, const uint8_t *labels,
                   const char *const *set_symbols, const int64_t *set_symbol_offsets,
                   int64_t G, const char *metric, int64_t P, int sort_descending,
                   int swap_labels, const char *dump_prefix, void *stats_out,
                   gsea_timings *t) {
    const int mid = gsea_metric_id(metric);
    if (mid < 0) return proxy_fail(GSEA_ERR_UNKNOWN_METRIC, "no known metric specified");
    if (!exprs || !gene_symbols || !labels || !set_symbol_offsets || !stats_out)
        return proxy_fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    SetCSR csr = map_sets_impl(
        N, [&](int64_t g) { return std::string(gene_symbols[g]); }, G,
        set_symbol_offsets, [&](int64_t i) { return std::string(set_symbols[i]); });
    const int flags = (sort_descending ? 0 : GSEA_FLAG_ASCENDING) |
                      (swap_labels ? GSEA_FLAG_SWAP_LABELS : 0);
    const std::string dump = dump_prefix ? dump_prefix : "";
    if (double_precision)
        return proxy_run_t<double>((const double *)exprs, N, S, labels, csr, G, mid,
                                   P, flags, dump, (double *)stats_out, t);
    return proxy_run_t<float>((const float *)exprs, N, S, labels, csr, G, mid, P,
                              flags, dump, (float *)stats_out, t);
}

}  // extern "C"

// ---- C++ proxy ABI (cudaGSEA_wrapper.cuh:31-61) ----

std::vector<proxy_strng_t> listCudaDevices_proxy() {
    std::vector<proxy_strng_t> out;
    int n = 0;
    if (gsea_device_count(&n)) return out;
    for (int i = 0; i < n; ++i) {
        char name[256] = {0};
        size_t bytes = 0;
        gsea_device_name(i, name, sizeof(name), &bytes);
        // cudaGSEA_wrapper.cu:87-104: "device <id>: <name> with <GiB> GiB RAM",
        // the memory size rounded up to whole GiB
        const size_t gib = (bytes + ((size_t)1 << 30) - 1) >> 30;
        out.push_back("device " + std::to_string(i) + ": " + name + " with " +
                      std::to_string(gib) + " GiB RAM");
    }
    return out;
}

int getCudaDevice_proxy() {
    int d = 0;
    gsea_get_device(&d);
    return d;
}

// cudaGSEA_wrapper.cu:126-137: an id that cannot be selected leaves the
// current device active and prints the reference's fallback warning
int setCudaDevice_proxy(int device) {
    int n = 0;
    gsea_device_count(&n);
    if (device >= 0 && device < n) gsea_set_device(device);
    const int active = getCudaDevice_proxy();
    if (active != device)
        std::cout << "WARNING: invalid device id, fallback to device " << active << "."
                  << std::endl;
    return active;
}

std::vector<float> gsea_single_proxy(std::vector<float> &exprs,
                                     std::vector<std::string> &gsymb,
                                     std::vector<std::string> &plist,
                                     std::vector<char> &labels,
                                     std::vector<std::string> &pname,
                                     std::vector<std::vector<std::string>> &pathw,
                                     std::string metric, size_t num_perms,
                                     bool sort_direction, bool swap_labels,
                                     std::string dump_filename) {
    return cxx_proxy<float>(exprs, gsymb, plist, labels, pname, pathw, metric,
                            num_perms, sort_direction, swap_labels, dump_filename);
}

std::vector<double> gsea_double_proxy(std::vector<double> &exprs,
                                      std::vector<std::string> &gsymb,
                                      std::vector<std::string> &plist,
                                      std::vector<char> &labels,
                                      std::vector<std::string> &pname,
                                      std::vector<std::vector<std::string>> &pathw,
                                      std::string metric, size_t num_perms,
                                      bool sort_direction, bool swap_labels,
                                      std::string dump_filename) {
    return cxx_proxy<double>(exprs, gsymb, plist, labels, pname, pathw, metric,
                             num_perms, sort_direction, swap_labels, dump_filename);
}
