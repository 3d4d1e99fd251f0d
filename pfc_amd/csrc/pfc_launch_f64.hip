// pfc_launch_f64.hip -- instantiates every double kernel and launcher.
#include "pfc_launch_impl.hpp"

template struct pfc::Launch<double>;
