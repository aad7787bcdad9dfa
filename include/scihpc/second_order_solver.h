// SciHPC's scihpc/second_order_solver.h for drivers built against the MI355X library: the
// declarations of that header, with the same names and signatures, live in
// include/scihpc_facade.hpp (namespace scihpc_gpu, brought into scope here).
#pragma once
#include "../scihpc_facade.hpp"
using namespace scihpc_gpu;
