// k_metric for fp64
#define GSEA_T double
#include "launch_metric.inc"
