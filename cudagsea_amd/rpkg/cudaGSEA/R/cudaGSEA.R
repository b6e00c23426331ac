# R entry points of the cudaGSEA package; each one is a .Call into
# src/gsea_rcpp.cpp.  Names, arguments and defaults follow the reference
# package (cudaGSEA/R/GSEA.R, cudaGSEA/R/DataReader.R, cudaGSEA/R/CUDA.R).

# Enrichment of each gene set against the phenotype labels, with
# numPermutations label permutations (the first one is the identity).
# exprsData: data.frame, genes x samples, row names = gene symbols.
# labelList: 0/1 per sample.  geneSets: named list of gene symbol vectors.
# Returns a data.frame with columns ES, NES, NP, FWER, FDR and one row per
# set, or FALSE when checkInput finds invalid data.
GSEA <- function(exprsData, labelList, geneSets, numPermutations,
                 metricString = "", dumpFileName = "", checkInput = TRUE,
                 doublePrecision = FALSE) {
    .Call("GSEA", exprsData, labelList, geneSets, numPermutations,
          metricString, dumpFileName, checkInput, doublePrecision,
          PACKAGE = "cudaGSEA")
}

# Broad file formats
loadLabelsFromCLS <- function(clsFileName) {
    .Call("loadLabelsFromCLS", clsFileName, PACKAGE = "cudaGSEA")
}

loadGeneSetsFromGMT <- function(gmtFileName) {
    .Call("loadGeneSetsFromGMT", gmtFileName, PACKAGE = "cudaGSEA")
}

loadExpressionDataFromGCT <- function(gctFileName) {
    .Call("loadExpressionDataFromGCT", gctFileName, PACKAGE = "cudaGSEA")
}

# GPUs: "device <id>: <name> with <GiB> GiB RAM" per device, the active
# device, and selection (an invalid id keeps the active device and warns)
listCudaDevices <- function() {
    .Call("listCudaDevices", PACKAGE = "cudaGSEA")
}

getCudaDevice <- function() {
    .Call("getCudaDevice", PACKAGE = "cudaGSEA")
}

setCudaDevice <- function(deviceId) {
    .Call("setCudaDevice", deviceId, PACKAGE = "cudaGSEA")
}
