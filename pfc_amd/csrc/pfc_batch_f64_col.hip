// pfc_batch_f64_col.hip -- instantiates the double batch col kernels and launchers.
#define PFC_BATCH_INSTANTIATE 1
#include "pfc_batch.hpp"

template void pfc::BatchLaunch<double>::col(int, bool, const pfc::BatchItem*, int,
                                         const pfc::BatchSys<double>*, hipStream_t);
template int pfc::BatchLaunch<double>::col_scratch(int);
