// k_metric for fp32
#define GSEA_T float
#include "launch_metric.inc"
