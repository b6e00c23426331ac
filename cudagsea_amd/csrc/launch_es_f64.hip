// k_es for fp64
#define GSEA_T double
#include "launch_es.inc"
