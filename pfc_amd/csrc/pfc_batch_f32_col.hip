// pfc_batch_f32_col.hip -- instantiates the float batch col kernels and launchers.
#define PFC_BATCH_INSTANTIATE 1
#include "pfc_batch.hpp"

template void pfc::BatchLaunch<float>::col(int, bool, const pfc::BatchItem*, int,
                                         const pfc::BatchSys<float>*, hipStream_t);
template int pfc::BatchLaunch<float>::col_scratch(int);
