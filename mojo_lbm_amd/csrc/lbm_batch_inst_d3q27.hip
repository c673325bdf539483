// Batched-case launchers of the Q=27 velocity set, all storages.
#include "lbm_batch_impl.hpp"

namespace lbm {
template struct BatchLaunch<Lattice<27>, StoreF32>;
template struct BatchLaunch<Lattice<27>, StoreF16C>;
template struct BatchLaunch<Lattice<27>, StoreF64>;
}  // namespace lbm
