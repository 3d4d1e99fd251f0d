// pfc_batch_f64_row.hip -- instantiates the double batch row kernels and launchers.
#define PFC_BATCH_INSTANTIATE 1
#include "pfc_batch.hpp"

template void pfc::BatchLaunch<double>::row(int, bool, const pfc::BatchItem*, int,
                                         const pfc::BatchSys<double>*, const int*, int,
                                         hipStream_t);
template int pfc::BatchLaunch<double>::row_scratch(int);
