// Compile-only check of include/powergrid/powergrid.hpp: every class template
// of the reference-shaped API instantiated explicitly, members included, so
// that a template no driver uses still has to compile.
#include "powergrid/powergrid.hpp"

namespace powergrid {
template class Gnufft<float>;
template class Gdft<float>;
template class TimeSegmentation<float, Gnufft<float>>;
template class SENSE<float, TimeSegmentation<float, Gnufft<float>>>;
template class SENSE<float, Gdft<float>>;
template class pcSenseTimeSeg<float>;
template class mpipcSENSETimeSeg<float>;
template class QuadPenalty<float>;
template class TVPenalty<float>;
}  // namespace powergrid

int main() { return 0; }
