// placeholder definition until unet.hip lands
#include "learned.hpp"
namespace scihpc {
void LearnedHook::apply(const double *, double *, hipStream_t) {}
}
