// pfc_launch_f32.hip -- instantiates every float kernel and launcher.
#include "pfc_launch_impl.hpp"

template struct pfc::Launch<float>;
