// pfc_batch_f32_row.hip -- instantiates the float batch row kernels and launchers.
#define PFC_BATCH_INSTANTIATE 1
#include "pfc_batch.hpp"

template void pfc::BatchLaunch<float>::row(int, bool, const pfc::BatchItem*, int,
                                         const pfc::BatchSys<float>*, const int*, int,
                                         hipStream_t);
template int pfc::BatchLaunch<float>::row_scratch(int);
