// k_es for fp32
#define GSEA_T float
#include "launch_es.inc"
