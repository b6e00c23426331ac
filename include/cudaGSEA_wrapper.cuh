// cudaGSEA_wrapper.cuh -- C++ proxy ABI expected by the reference's Rcpp glue
// (cudaGSEA/src/cudaGSEA_wrapper.cpp includes a header of this name).  The
// declarations must match cudaGSEA/src/cudaGSEA_wrapper.cuh:6-61 exactly so
// that the glue compiles and links against libgsea_hip.so unchanged; the
// implementations live in cudagsea_amd/csrc/gsea_proxy.cpp (GSEA entry points
// and device management) and cudagsea_amd/csrc/gsea_io.cpp (GCT/GMT/CLS).
#ifndef CUDAGSEA_AMD_WRAPPER_CUH
#define CUDAGSEA_AMD_WRAPPER_CUH

#include <cstddef>
#include <iostream>
#include <string>
#include <vector>

// value types exchanged with R (cudaGSEA_wrapper.cuh:6-9)
typedef std::string proxy_strng_t;
typedef double      proxy_value_t;
typedef size_t      proxy_label_t;
typedef size_t      proxy_index_t;

// Broad file formats (cudaGSEA_wrapper.cuh:15-25)
void loadLabelsFromCLS_proxy(std::string, std::vector<proxy_label_t> &);
void loadGeneSetsFromGMT_proxy(std::string, std::vector<proxy_strng_t> &,
                               std::vector<std::vector<proxy_strng_t>> &);
void loadExpressionDataFromGCT_proxy(std::string, std::vector<proxy_value_t> &,
                                     std::vector<proxy_strng_t> &,
                                     std::vector<proxy_strng_t> &);

// device information and selection (cudaGSEA_wrapper.cuh:31-33)
std::vector<proxy_strng_t> listCudaDevices_proxy();
int getCudaDevice_proxy();
int setCudaDevice_proxy(int);

// GSEA (cudaGSEA_wrapper.cuh:39-61).  Arguments: expression data sample-major
// [S][N], gene symbols (N), sample names (S), labels (S, 0/1), set names (G),
// set symbol lists (G), metric name, permutations, sort descending, swap
// labels, dump prefix.  Returns [ES|NES|NP|FWER|FDR], each of length G.
std::vector<float> gsea_single_proxy(std::vector<float> &,
                                     std::vector<std::string> &,
                                     std::vector<std::string> &,
                                     std::vector<char> &,
                                     std::vector<std::string> &,
                                     std::vector<std::vector<std::string>> &,
                                     std::string, size_t, bool, bool,
                                     std::string);

std::vector<double> gsea_double_proxy(std::vector<double> &,
                                      std::vector<std::string> &,
                                      std::vector<std::string> &,
                                      std::vector<char> &,
                                      std::vector<std::string> &,
                                      std::vector<std::vector<std::string>> &,
                                      std::string, size_t, bool, bool,
                                      std::string);

#endif
