// Batched-case launchers of the Q=19 velocity set, all storages.
#include "lbm_batch_impl.hpp"

namespace lbm {
template struct BatchLaunch<Lattice<19>, StoreF32>;
template struct BatchLaunch<Lattice<19>, StoreF16C>;
template struct BatchLaunch<Lattice<19>, StoreF64>;
}  // namespace lbm
