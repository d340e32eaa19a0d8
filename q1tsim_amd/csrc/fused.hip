// Fused complex64 path: dense k-qubit blocks applied with FP32 MFMA.
#include "sweep.h"
