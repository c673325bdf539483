// Kernel instantiations for one (velocity set, storage) pair; split per file so
// the library builds in parallel.
#include "lbm_launch_impl.hpp"

namespace lbm {
template struct Launch<Lattice<9>, StoreF16C>;
}  // namespace lbm
