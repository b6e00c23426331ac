// gsea_error.hpp -- exception carrying a cudaGSEA error code
// (cudaGSEA/src/include/error_codes.cuh), thrown by the file loaders where the
// reference calls exit(code); the command line turns it back into that exit
// status, Rcpp into an R error.
#pragma once

#include <stdexcept>
#include <string>

namespace gsea {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

}  // namespace gsea
