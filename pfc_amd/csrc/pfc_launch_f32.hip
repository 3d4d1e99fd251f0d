// pfc_launch_f32.hip -- instantiates every float kernel and launcher.
#include "pfc_launch.hpp"

template struct pfc::Launch<float>;
