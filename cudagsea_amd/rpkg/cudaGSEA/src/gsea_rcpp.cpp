// .Call entry points of the cudaGSEA R package on the C++ proxy ABI of
// libgsea_hip.so (include/cudaGSEA_wrapper.cuh).  The R-visible behaviour is
// that of the reference package's glue, cudaGSEA/src/cudaGSEA_wrapper.cpp:
// the same entry point names and arguments, the same input checks with the
// same messages (an invalid input prints one "INVALID DATA" line and returns
// FALSE), the same result rows and row names.  Differences: the result has
// a fifth column, FDR (the reference's engine leaves it NaN and its glue
// drops it), and errors from the engine (an unknown metric, a GPU failure)
// become R errors instead of ending the R session.
#include <Rcpp.h>

#include <cmath>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "cudaGSEA_wrapper.cuh"

namespace {

SEXP invalid(const char *msg) {
    std::cout << "INVALID DATA: " << msg << std::endl;
    return Rcpp::wrap(false);
}

// the proxies return [ES | NES | NP | FWER | FDR], each G long
template <class T>
SEXP result_frame(const std::vector<T> &r, size_t G, const std::vector<std::string> &names) {
    auto column = [&](size_t k) {
        std::vector<double> v(r.begin() + k * G, r.begin() + (k + 1) * G);
        return Rcpp::NumericVector(v.begin(), v.end());
    };
    Rcpp::DataFrame frame = Rcpp::DataFrame::create(
        Rcpp::Named("ES") = column(0), Rcpp::Named("NES") = column(1),
        Rcpp::Named("NP") = column(2), Rcpp::Named("FWER") = column(3),
        Rcpp::Named("FDR") = column(4));
    frame.attr("row.names") = Rcpp::CharacterVector(names.begin(), names.end());
    return frame;
}

}  // namespace

// ---------------------------------------------------------------------------
// Broad file formats
// ---------------------------------------------------------------------------
RcppExport SEXP loadLabelsFromCLS(SEXP clsFileName) {
    BEGIN_RCPP
    std::vector<proxy_label_t> labels;
    loadLabelsFromCLS_proxy(Rcpp::as<std::string>(clsFileName), labels);
    return Rcpp::wrap(labels);
    END_RCPP
}

RcppExport SEXP loadGeneSetsFromGMT(SEXP gmtFileName) {
    BEGIN_RCPP
    std::vector<proxy_strng_t> names;
    std::vector<std::vector<proxy_strng_t>> sets;
    loadGeneSetsFromGMT_proxy(Rcpp::as<std::string>(gmtFileName), names, sets);
    Rcpp::List out(sets.size());
    for (size_t i = 0; i < sets.size(); ++i) out[i] = Rcpp::wrap(sets[i]);
    out.attr("names") = Rcpp::wrap(names);
    return out;
    END_RCPP
}

RcppExport SEXP loadExpressionDataFromGCT(SEXP gctFileName) {
    BEGIN_RCPP
    std::vector<proxy_value_t> values;   // gene-major [N][S]
    std::vector<proxy_strng_t> genes, samples;
    loadExpressionDataFromGCT_proxy(Rcpp::as<std::string>(gctFileName), values, genes,
                                    samples);
    const size_t N = genes.size(), S = samples.size();
    Rcpp::NumericMatrix m((int)N, (int)S);   // R matrices are column-major
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < S; ++j) m((int)i, (int)j) = values[i * S + j];
    m.attr("dimnames") = Rcpp::List::create(Rcpp::wrap(genes), Rcpp::wrap(samples));
    Rcpp::Function as_data_frame("as.data.frame");
    return as_data_frame(m);
    END_RCPP
}

// ---------------------------------------------------------------------------
// GPUs
// ---------------------------------------------------------------------------
RcppExport SEXP listCudaDevices() {
    BEGIN_RCPP
    return Rcpp::wrap(listCudaDevices_proxy());
    END_RCPP
}

RcppExport SEXP getCudaDevice() {
    BEGIN_RCPP
    return Rcpp::wrap(getCudaDevice_proxy());
    END_RCPP
}

RcppExport SEXP setCudaDevice(SEXP deviceId) {
    BEGIN_RCPP
    return Rcpp::wrap(setCudaDevice_proxy(Rcpp::as<int>(deviceId)));
    END_RCPP
}

// ---------------------------------------------------------------------------
// GSEA
// ---------------------------------------------------------------------------
RcppExport SEXP GSEA(SEXP exprsData, SEXP labelList, SEXP geneSets, SEXP numPermutations,
                     SEXP metricString, SEXP dumpFileName, SEXP checkInput,
                     SEXP doublePrecision) {
    BEGIN_RCPP
    try {
        // expression table: genes x samples, gene symbols as row names
        Rcpp::DataFrame frame(exprsData);
        Rcpp::CharacterVector col_names = frame.attr("names");
        Rcpp::CharacterVector row_names = frame.attr("row.names");
        const std::vector<std::string> samples =
            Rcpp::as<std::vector<std::string>>(col_names);
        std::vector<std::string> genes = Rcpp::as<std::vector<std::string>>(row_names);
        Rcpp::Function as_matrix("as.matrix");
        Rcpp::NumericMatrix table = as_matrix(frame);
        const size_t rows = (size_t)table.nrow(), cols = (size_t)table.ncol();

        // gene sets and their names ("1".."n" when the list is not named)
        Rcpp::List set_list(geneSets);
        std::vector<std::vector<std::string>> sets;
        sets.reserve((size_t)set_list.size());
        for (R_xlen_t i = 0; i < set_list.size(); ++i)
            sets.push_back(Rcpp::as<std::vector<std::string>>(set_list[i]));
        std::vector<std::string> names;
        SEXP set_names = set_list.attr("names");
        if (set_names != R_NilValue) names = Rcpp::as<std::vector<std::string>>(set_names);
        bool enumerate = names.size() != sets.size();
        for (const std::string &n : names) enumerate |= n == "NA";
        if (enumerate) {
            names.resize(sets.size());
            for (size_t j = 0; j < sets.size(); ++j) names[j] = std::to_string(j + 1);
            std::cout << "WARNING: We enumerated your gene set names from 1 to n."
                      << std::endl;
        }

        const std::vector<proxy_label_t> label_values =
            Rcpp::as<std::vector<proxy_label_t>>(labelList);
        std::vector<char> labels(label_values.begin(), label_values.end());
        const proxy_index_t num_perms = Rcpp::as<proxy_index_t>(numPermutations);
        const std::string metric = Rcpp::as<std::string>(metricString);
        const std::string dump = Rcpp::as<std::string>(dumpFileName);
        const bool double_precision = Rcpp::as<bool>(doublePrecision);

        if (Rcpp::as<bool>(checkInput)) {
            if (rows != genes.size())
                return invalid("the vertical dimension and the number of row names do "
                               "not comply.");
            if (cols != samples.size())
                return invalid("the horizontal dimension and the number of column names "
                               "do not comply.");
            if (label_values.size() != cols)
                return invalid("Your label list has a different amount of entries as the "
                               "number of columns in your gene  expression data frame.");
            for (proxy_label_t v : label_values)
                if (v != 0 && v != 1)
                    return invalid("Label vector may only contain zeros and ones since "
                                   "cudaGSEA only supports two classes of phenotypes.");
            double lowest = std::numeric_limits<double>::infinity();
            for (double v : table) {
                if (std::isnan(v))
                    return invalid("expression table has NaN entries that cannot be "
                                   "handled by the GSEA algorithm.");
                lowest = v < lowest ? v : lowest;
            }
            if (lowest < 0)
                std::cout << "WARNING: There are negative expression values in the "
                             "provided table. Some metrics cannot handle negative "
                             "amplitudes e.g. the log2_ratio_of_classes family."
                          << std::endl;
            if (sets.empty()) return invalid("you must at least specify one gene set.");
            for (const auto &s : sets)
                if (s.empty())
                    return invalid("There is an empty gene set (no gene identifiers) in "
                                   "your list of gene sets.");
        }

        // the column-major genes x samples matrix is the sample-major [S][N]
        // block the proxies take; descending order, labels as given
        std::vector<std::string> sample_names = samples;
        if (double_precision) {
            std::vector<double> x(table.begin(), table.end());
            const std::vector<double> r = gsea_double_proxy(
                x, genes, sample_names, labels, names, sets, metric, num_perms, true, false,
                dump);
            return result_frame(r, sets.size(), names);
        }
        std::vector<float> x(table.begin(), table.end());
        const std::vector<float> r = gsea_single_proxy(
            x, genes, sample_names, labels, names, sets, metric, num_perms, true, false, dump);
        return result_frame(r, sets.size(), names);
    } catch (Rcpp::not_compatible &) {
        return invalid("cannot cast data. Make sure you fill in the correct datastructures.");
    }
    END_RCPP
}
